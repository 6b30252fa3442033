"""Python front end of the MI355X JPEG XS library (libsvtjpegxs_amd.so).

Mirrors the reference C API (Source/API/SvtJpegxsEnc.h, SvtJpegxsDec.h):
the ctypes structures below have the reference's layout, and the Encoder /
Decoder classes call svt_jpeg_xs_encoder_* / svt_jpeg_xs_decoder_* exactly as
the reference's own tools do (load defaults, init, send, get, close).  The
Plan class drives the device-resident batch entry points of svt_jpegxs_amd.h,
and the Server class its serving loop.

The package has no CPU fallback: importing it on a machine without the built
library raises, and the GPU entry points fail loudly without a GPU.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsvtjpegxs_amd.so")

API_MAJOR, API_MINOR = 0, 10

# Stages timed by Plan.run_timed (SVT_JPEG_XS_AMD_STAGES, include/svt_jpegxs_amd.h)
STAGE_NAMES = ["dwt1", "dwt2", "rate", "pack"]
STAGES = len(STAGE_NAMES)

# SvtJxsErrorType_t values (Source/API/SvtJpegxs.h:130-159)
ERR_NONE = 0
ERR_BAD_PARAMETER = -0x7FFFEFFB          # 0x80001005
ERR_INSUFFICIENT_RESOURCES = -0x7FFFF000  # 0x80001000
ERR_EMPTY_QUEUE = -0x7FFFDFCD            # 0x80002033
ERR_ENCODE_FRAME = -0x7FFFDFCB           # 0x80002035
ERR_DEC_INVALID_POINTER = -0x7FFFD000    # 0x80003000
ERR_DEC_INVALID_BITSTREAM = -0x7FFFCFFF  # 0x80003001
ERR_DEC_TOO_SHORT = -0x7FFFCFFD          # 0x80003003
ERR_DEC_CONFIG_CHANGE = -0x7FFFCFFC      # 0x80003004
ERR_DEC_EOC = -0x7FFFCFFB                # 0x80003005
ERR_UNDEFINED = -0x7FFFEFFF              # 0x80001001

FORMAT_YUV420, FORMAT_YUV422, FORMAT_YUV444, FORMAT_PACKED444 = 2, 3, 4, 21

_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_void_p)


class EncoderApi(ctypes.Structure):
    """svt_jpeg_xs_encoder_api_t (Source/API/SvtJpegxsEnc.h:18-93)."""
    _fields_ = [
        ("source_width", ctypes.c_uint32), ("source_height", ctypes.c_uint32),
        ("input_bit_depth", ctypes.c_uint8), ("colour_format", ctypes.c_int32),
        ("bpp_numerator", ctypes.c_uint32), ("bpp_denominator", ctypes.c_uint32),
        ("ndecomp_v", ctypes.c_uint32), ("ndecomp_h", ctypes.c_uint32),
        ("quantization", ctypes.c_uint32), ("slice_height", ctypes.c_uint32),
        ("use_cpu_flags", ctypes.c_uint64), ("threads_num", ctypes.c_uint32),
        ("cpu_profile", ctypes.c_uint8), ("print_bands_info", ctypes.c_uint32),
        ("coding_signs_handling", ctypes.c_uint32), ("coding_significance", ctypes.c_uint32),
        ("rate_control_mode", ctypes.c_uint32), ("coding_vertical_prediction_mode", ctypes.c_uint32),
        ("verbose", ctypes.c_uint32),
        ("callback_send_data_available", _CB), ("callback_send_data_available_context", ctypes.c_void_p),
        ("callback_get_data_available", _CB), ("callback_get_data_available_context", ctypes.c_void_p),
        ("slice_packetization_mode", ctypes.c_uint8), ("private_ptr", ctypes.c_void_p),
        ("padding", ctypes.c_uint8 * 64)]


class DecoderApi(ctypes.Structure):
    """svt_jpeg_xs_decoder_api_t (Source/API/SvtJpegxsDec.h:16-57)."""
    _fields_ = [
        ("use_cpu_flags", ctypes.c_uint64), ("verbose", ctypes.c_uint32), ("threads_num", ctypes.c_uint32),
        ("packetization_mode", ctypes.c_uint8), ("proxy_mode", ctypes.c_int32),
        ("callback_send_data_available", _CB), ("callback_send_data_available_context", ctypes.c_void_p),
        ("callback_get_data_available", _CB), ("callback_get_data_available_context", ctypes.c_void_p),
        ("private_ptr", ctypes.c_void_p), ("padding", ctypes.c_uint8 * 64)]


class _CompCfg(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("byte_size", ctypes.c_uint32)]


class ImageConfig(ctypes.Structure):
    """svt_jpeg_xs_image_config_t (Source/API/SvtJpegxs.h:73-84)."""
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("bit_depth", ctypes.c_uint8),
                ("format", ctypes.c_int32), ("components_num", ctypes.c_uint8), ("components", _CompCfg * 4)]


class ImageBuffer(ctypes.Structure):
    _fields_ = [("data_yuv", ctypes.c_void_p * 4), ("stride", ctypes.c_uint32 * 4), ("alloc_size", ctypes.c_uint32 * 4),
                ("release_ctx_ptr", ctypes.c_void_p), ("ready_to_release", ctypes.c_uint8)]


class BitstreamBuffer(ctypes.Structure):
    _fields_ = [("buffer", ctypes.c_void_p), ("allocation_size", ctypes.c_uint32), ("used_size", ctypes.c_uint32),
                ("release_ctx_ptr", ctypes.c_void_p), ("ready_to_release", ctypes.c_uint8),
                ("last_packet_in_frame", ctypes.c_uint8)]


class Frame(ctypes.Structure):
    """svt_jpeg_xs_frame_t (Source/API/SvtJpegxs.h:121-128)."""
    _fields_ = [("image", ImageBuffer), ("bitstream", BitstreamBuffer), ("user_prv_ctx_ptr", ctypes.c_void_p)]


class FrameSpec(ctypes.Structure):
    """svt_jpeg_xs_amd_frame_spec_t (include/svt_jpegxs_amd.h)."""
    _fields_ = [("encoder", ctypes.POINTER(EncoderApi)), ("d_planes", ctypes.c_void_p * 4),
                ("strides", ctypes.c_uint32 * 4), ("d_out", ctypes.c_void_p)]


class Tile(ctypes.Structure):
    """svt_jpeg_xs_amd_tile_t (include/svt_jpegxs_amd.h)."""
    _fields_ = [("width", ctypes.c_uint32), ("levels", ctypes.c_uint32), ("vertical", ctypes.c_uint32),
                ("chunk", ctypes.c_uint32), ("frame_us", ctypes.c_float), ("stage", ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    """svt_jpeg_xs_amd_server_config_t (include/svt_jpegxs_amd.h)."""
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("max_batch", ctypes.c_uint32), ("max_inflight", ctypes.c_uint32)]


class Completion(ctypes.Structure):
    """svt_jpeg_xs_amd_completion_t (include/svt_jpegxs_amd.h)."""
    _fields_ = [("tag", ctypes.c_uint64), ("status", ctypes.c_int32), ("batch_frames", ctypes.c_uint32),
                ("submit_ns", ctypes.c_uint64), ("launch_ns", ctypes.c_uint64), ("done_ns", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    """svt_jpeg_xs_amd_server_stats_t (include/svt_jpegxs_amd.h)."""
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("largest_batch", ctypes.c_uint64), ("hbm_budget", ctypes.c_uint64), ("hbm_in_use", ctypes.c_uint64),
                ("hbm_peak", ctypes.c_uint64), ("queued", ctypes.c_uint32), ("inflight", ctypes.c_uint32)]


class DecodeSpec(ctypes.Structure):
    """svt_jpeg_xs_amd_decode_spec_t (include/svt_jpegxs_amd.h)."""
    _fields_ = [("decoder", ctypes.POINTER(DecoderApi)), ("d_codestream", ctypes.c_void_p),
                ("codestream_size", ctypes.c_uint32), ("d_planes", ctypes.c_void_p * 4),
                ("strides", ctypes.c_uint32 * 4), ("d_precinct_offsets", ctypes.c_void_p),
                ("precinct_offsets_stride", ctypes.c_uint32)]


# Every C-ABI symbol the library exports (include/*.h).
EXPORTED_SYMBOLS = [
    "svt_jpeg_xs_encoder_load_default_parameters", "svt_jpeg_xs_encoder_get_image_config",
    "svt_jpeg_xs_encoder_init", "svt_jpeg_xs_encoder_close", "svt_jpeg_xs_encoder_send_picture",
    "svt_jpeg_xs_encoder_get_packet",
    "svt_jpeg_xs_decoder_init", "svt_jpeg_xs_decoder_close", "svt_jpeg_xs_decoder_get_single_frame_size",
    "svt_jpeg_xs_decoder_get_single_frame_size_with_proxy", "svt_jpeg_xs_decoder_send_frame",
    "svt_jpeg_xs_decoder_send_packet", "svt_jpeg_xs_decoder_get_frame", "svt_jpeg_xs_decoder_send_eoc",
    "svt_jpeg_xs_image_buffer_alloc", "svt_jpeg_xs_image_buffer_free", "svt_jpeg_xs_bitstream_alloc",
    "svt_jpeg_xs_bitstream_free", "svt_jpeg_xs_frame_pool_alloc", "svt_jpeg_xs_frame_pool_free",
    "svt_jpeg_xs_frame_pool_get", "svt_jpeg_xs_frame_pool_release",
    "svt_jpeg_xs_amd_plan_create", "svt_jpeg_xs_amd_plan_run", "svt_jpeg_xs_amd_plan_run_timed",
    "svt_jpeg_xs_amd_plan_status", "svt_jpeg_xs_amd_plan_destroy", "svt_jpeg_xs_amd_plan_footprint",
    "svt_jpeg_xs_amd_version",
    "svt_jpeg_xs_amd_decode_plan_create", "svt_jpeg_xs_amd_decode_plan_run", "svt_jpeg_xs_amd_decode_plan_status",
    "svt_jpeg_xs_amd_decode_plan_destroy", "svt_jpeg_xs_amd_plan_precinct_offsets",
    "svt_jpeg_xs_amd_server_create", "svt_jpeg_xs_amd_server_submit", "svt_jpeg_xs_amd_server_poll",
    "svt_jpeg_xs_amd_server_stats", "svt_jpeg_xs_amd_server_now_ns", "svt_jpeg_xs_amd_server_destroy",
    "svt_jpeg_xs_amd_autotune", "svt_jpeg_xs_amd_encode_noalloc",
]

_lib = None


def lib():
    """Loads the library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsvtjpegxs_amd.so is not built: run svt-jpeg-xs_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        E = ctypes.POINTER(EncoderApi)
        D = ctypes.POINTER(DecoderApi)
        L.svt_jpeg_xs_encoder_load_default_parameters.argtypes = [ctypes.c_uint64, ctypes.c_uint64, E]
        L.svt_jpeg_xs_encoder_get_image_config.argtypes = [ctypes.c_uint64, ctypes.c_uint64, E,
                                                           ctypes.POINTER(ImageConfig), ctypes.POINTER(ctypes.c_uint32)]
        L.svt_jpeg_xs_encoder_init.argtypes = [ctypes.c_uint64, ctypes.c_uint64, E]
        L.svt_jpeg_xs_encoder_close.argtypes = [E]
        L.svt_jpeg_xs_encoder_close.restype = None
        L.svt_jpeg_xs_encoder_send_picture.argtypes = [E, ctypes.POINTER(Frame), ctypes.c_uint8]
        L.svt_jpeg_xs_encoder_get_packet.argtypes = [E, ctypes.POINTER(Frame), ctypes.c_uint8]
        L.svt_jpeg_xs_decoder_init.argtypes = [ctypes.c_uint64, ctypes.c_uint64, D, ctypes.c_void_p, ctypes.c_size_t,
                                               ctypes.POINTER(ImageConfig)]
        L.svt_jpeg_xs_decoder_close.argtypes = [D]
        L.svt_jpeg_xs_decoder_close.restype = None
        L.svt_jpeg_xs_decoder_get_single_frame_size_with_proxy.argtypes = [
            ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ImageConfig), ctypes.POINTER(ctypes.c_uint32),
            ctypes.c_uint32, ctypes.c_int32]
        L.svt_jpeg_xs_decoder_get_single_frame_size.argtypes = [
            ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ImageConfig), ctypes.POINTER(ctypes.c_uint32),
            ctypes.c_uint32]
        L.svt_jpeg_xs_decoder_send_frame.argtypes = [D, ctypes.POINTER(Frame), ctypes.c_uint8]
        L.svt_jpeg_xs_decoder_send_packet.argtypes = [D, ctypes.POINTER(Frame), ctypes.POINTER(ctypes.c_uint32)]
        L.svt_jpeg_xs_decoder_get_frame.argtypes = [D, ctypes.POINTER(Frame), ctypes.c_uint8]
        L.svt_jpeg_xs_decoder_send_eoc.argtypes = [D]
        L.svt_jpeg_xs_amd_decode_plan_create.argtypes = [ctypes.POINTER(DecodeSpec), ctypes.c_uint32]
        L.svt_jpeg_xs_amd_decode_plan_create.restype = ctypes.c_void_p
        L.svt_jpeg_xs_amd_decode_plan_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.svt_jpeg_xs_amd_decode_plan_status.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
        L.svt_jpeg_xs_amd_decode_plan_destroy.argtypes = [ctypes.c_void_p]
        L.svt_jpeg_xs_amd_decode_plan_destroy.restype = None
        L.svt_jpeg_xs_amd_plan_precinct_offsets.argtypes = [ctypes.c_void_p, ctypes.c_uint32,
                                                            ctypes.POINTER(ctypes.c_void_p),
                                                            ctypes.POINTER(ctypes.c_uint32)]
        L.svt_jpeg_xs_amd_plan_create.argtypes = [ctypes.POINTER(FrameSpec), ctypes.c_uint32]
        L.svt_jpeg_xs_amd_plan_create.restype = ctypes.c_void_p
        L.svt_jpeg_xs_amd_plan_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
        L.svt_jpeg_xs_amd_plan_run_timed.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
        L.svt_jpeg_xs_amd_plan_status.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
        L.svt_jpeg_xs_amd_plan_destroy.argtypes = [ctypes.c_void_p]
        L.svt_jpeg_xs_amd_plan_destroy.restype = None
        L.svt_jpeg_xs_amd_plan_footprint.argtypes = [ctypes.POINTER(FrameSpec), ctypes.c_uint32]
        L.svt_jpeg_xs_amd_plan_footprint.restype = ctypes.c_uint64
        L.svt_jpeg_xs_amd_version.restype = ctypes.c_char_p
        L.svt_jpeg_xs_amd_autotune.argtypes = [E, ctypes.c_void_p, ctypes.POINTER(Tile), ctypes.c_uint32]
        L.svt_jpeg_xs_amd_autotune.restype = ctypes.c_int32
        L.svt_jpeg_xs_amd_encode_noalloc.argtypes = [ctypes.POINTER(FrameSpec), ctypes.c_void_p, ctypes.c_void_p]
        L.svt_jpeg_xs_amd_encode_noalloc.restype = ctypes.c_int32
        L.svt_jpeg_xs_amd_server_create.argtypes = [ctypes.POINTER(ServerConfig)]
        L.svt_jpeg_xs_amd_server_create.restype = ctypes.c_void_p
        L.svt_jpeg_xs_amd_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(FrameSpec), ctypes.c_uint64]
        L.svt_jpeg_xs_amd_server_poll.argtypes = [ctypes.c_void_p, ctypes.POINTER(Completion), ctypes.c_uint32,
                                                  ctypes.c_uint32]
        L.svt_jpeg_xs_amd_server_poll.restype = ctypes.c_uint32
        L.svt_jpeg_xs_amd_server_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerStats)]
        L.svt_jpeg_xs_amd_server_now_ns.argtypes = []
        L.svt_jpeg_xs_amd_server_now_ns.restype = ctypes.c_uint64
        L.svt_jpeg_xs_amd_server_destroy.argtypes = [ctypes.c_void_p]
        L.svt_jpeg_xs_amd_server_destroy.restype = None
        L.svt_jpeg_xs_image_buffer_alloc.argtypes = [ctypes.POINTER(ImageConfig)]
        L.svt_jpeg_xs_image_buffer_alloc.restype = ctypes.POINTER(ImageBuffer)
        L.svt_jpeg_xs_image_buffer_free.argtypes = [ctypes.POINTER(ImageBuffer)]
        L.svt_jpeg_xs_image_buffer_free.restype = None
        L.svt_jpeg_xs_bitstream_alloc.argtypes = [ctypes.c_uint32]
        L.svt_jpeg_xs_bitstream_alloc.restype = ctypes.POINTER(BitstreamBuffer)
        L.svt_jpeg_xs_bitstream_free.argtypes = [ctypes.POINTER(BitstreamBuffer)]
        L.svt_jpeg_xs_bitstream_free.restype = None
        L.svt_jpeg_xs_frame_pool_alloc.argtypes = [ctypes.POINTER(ImageConfig), ctypes.c_uint32, ctypes.c_uint32]
        L.svt_jpeg_xs_frame_pool_alloc.restype = ctypes.c_void_p
        L.svt_jpeg_xs_frame_pool_free.argtypes = [ctypes.c_void_p]
        L.svt_jpeg_xs_frame_pool_free.restype = None
        L.svt_jpeg_xs_frame_pool_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(Frame), ctypes.c_uint8]
        L.svt_jpeg_xs_frame_pool_get.restype = ctypes.c_int32
        L.svt_jpeg_xs_frame_pool_release.argtypes = [ctypes.c_void_p, ctypes.POINTER(Frame)]
        L.svt_jpeg_xs_frame_pool_release.restype = None
        for name in ("svt_jpeg_xs_encoder_load_default_parameters", "svt_jpeg_xs_encoder_get_image_config",
                     "svt_jpeg_xs_encoder_init", "svt_jpeg_xs_encoder_send_picture", "svt_jpeg_xs_encoder_get_packet",
                     "svt_jpeg_xs_decoder_init", "svt_jpeg_xs_decoder_get_single_frame_size_with_proxy",
                     "svt_jpeg_xs_decoder_get_single_frame_size", "svt_jpeg_xs_decoder_send_frame",
                     "svt_jpeg_xs_decoder_send_packet", "svt_jpeg_xs_decoder_get_frame", "svt_jpeg_xs_decoder_send_eoc",
                     "svt_jpeg_xs_amd_plan_run", "svt_jpeg_xs_amd_plan_run_timed", "svt_jpeg_xs_amd_plan_status",
                     "svt_jpeg_xs_amd_decode_plan_run", "svt_jpeg_xs_amd_decode_plan_status",
                     "svt_jpeg_xs_amd_plan_precinct_offsets",
                     "svt_jpeg_xs_amd_server_submit", "svt_jpeg_xs_amd_server_stats"):
            getattr(L, name).restype = ctypes.c_int32
        _lib = L
    return _lib


def missing_symbols():
    L = lib()
    return [s for s in EXPORTED_SYMBOLS if not hasattr(L, s)]


class JxsError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s failed with 0x%08x" % (what, code & 0xFFFFFFFF))
        self.code = code


class Encoder:
    """One encoder handle: svt_jpeg_xs_encoder_load_default_parameters, then
    the given fields, then svt_jpeg_xs_encoder_init."""

    def __init__(self, width, height, depth, fmt, bpp_num, bpp_den=1, decomp_v=2, decomp_h=5, quantization=0,
                 slice_height=16, significance=1, verbose=1, slice_packetization=0, rate_control_mode=1,
                 callback_send=None, callback_get=None, vertical_prediction=0, signs_handling=0):
        L = lib()
        self.api = EncoderApi()
        rc = L.svt_jpeg_xs_encoder_load_default_parameters(API_MAJOR, API_MINOR, ctypes.byref(self.api))
        if rc:
            raise JxsError(rc, "load_default_parameters")
        a = self.api
        # callbacks are called from the library's threads (SvtJpegxsEnc.h:57-73); the
        # ctypes thunks must outlive the encoder
        self._cbs = [_CB(lambda api, ctx, f=f: f()) if f else None for f in (callback_send, callback_get)]
        if self._cbs[0]:
            a.callback_send_data_available = self._cbs[0]
        if self._cbs[1]:
            a.callback_get_data_available = self._cbs[1]
        a.coding_vertical_prediction_mode = vertical_prediction
        a.coding_signs_handling = signs_handling
        a.source_width, a.source_height, a.input_bit_depth, a.colour_format = width, height, depth, fmt
        a.bpp_numerator, a.bpp_denominator, a.ndecomp_v, a.ndecomp_h = bpp_num, bpp_den, decomp_v, decomp_h
        a.quantization, a.slice_height, a.coding_significance, a.verbose = quantization, slice_height, significance, verbose
        a.slice_packetization_mode = slice_packetization
        a.rate_control_mode = rate_control_mode
        self.config = ImageConfig()
        nb = ctypes.c_uint32(0)
        rc = L.svt_jpeg_xs_encoder_get_image_config(API_MAJOR, API_MINOR, ctypes.byref(a), ctypes.byref(self.config),
                                                    ctypes.byref(nb))
        if rc:
            raise JxsError(rc, "get_image_config")
        self.bytes_per_frame = nb.value
        rc = L.svt_jpeg_xs_encoder_init(API_MAJOR, API_MINOR, ctypes.byref(a))
        if rc:
            raise JxsError(rc, "encoder_init")
        self.fmt = fmt
        self.depth = depth

    def close(self):
        if self.api.private_ptr:
            lib().svt_jpeg_xs_encoder_close(ctypes.byref(self.api))

    def autotune(self, stream=0):
        """Times the transform of one frame of this configuration with every
        supported tile width and keeps the fastest per job shape
        (svt_jpeg_xs_amd_autotune).  Returns the choices as dicts."""
        out = (Tile * 16)()
        n = lib().svt_jpeg_xs_amd_autotune(ctypes.byref(self.api), stream, out, 16)
        if n < 0:
            raise RuntimeError("svt_jpeg_xs_amd_autotune failed")
        return [{f: getattr(out[i], f) for f, _ in Tile._fields_} for i in range(min(n, 16))]

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _frame(self, planes, out):
        f = Frame()
        pix = 1 if self.depth <= 8 else 2
        for c, p in enumerate(planes):
            assert p.flags["C_CONTIGUOUS"]
            f.image.data_yuv[c] = p.ctypes.data
            f.image.stride[c] = p.shape[1]
            f.image.alloc_size[c] = p.nbytes
        if self.fmt == FORMAT_PACKED444:
            for c in (1, 2):
                f.image.data_yuv[c] = f.image.data_yuv[0]
                f.image.stride[c] = f.image.stride[0]
                f.image.alloc_size[c] = f.image.alloc_size[0]
        f.bitstream.buffer = out.ctypes.data
        f.bitstream.allocation_size = out.nbytes
        del pix
        return f

    def send(self, planes, out, ctx=None, blocking=1):
        f = self._frame(planes, out)
        f.user_prv_ctx_ptr = ctx
        return lib().svt_jpeg_xs_encoder_send_picture(ctypes.byref(self.api), ctypes.byref(f), blocking)

    def get(self, blocking=1):
        f = Frame()
        rc = lib().svt_jpeg_xs_encoder_get_packet(ctypes.byref(self.api), ctypes.byref(f), blocking)
        return rc, f

    def encode(self, planes):
        """One frame through send_picture / get_packet; returns (rc, bytes)."""
        out = np.zeros(self.bytes_per_frame, dtype=np.uint8)
        rc = self.send(planes, out)
        if rc:
            return rc, b""
        rc, f = self.get()
        n = f.bitstream.used_size
        return rc, out[:n].tobytes()


class Decoder:
    """svt_jpeg_xs_decoder_init on the first codestream, then send/get."""

    def __init__(self, codestream, verbose=1, proxy_mode=0, packetization_mode=0, callback_send=None,
                 callback_get=None):
        L = lib()
        self.api = DecoderApi()
        self._cbs = [_CB(lambda api, ctx, f=f: f()) if f else None for f in (callback_send, callback_get)]
        if self._cbs[0]:
            self.api.callback_send_data_available = self._cbs[0]
        if self._cbs[1]:
            self.api.callback_get_data_available = self._cbs[1]
        self.api.verbose = verbose
        self.api.proxy_mode = proxy_mode
        self.api.packetization_mode = packetization_mode
        self.config = ImageConfig()
        buf = np.frombuffer(codestream, dtype=np.uint8)
        rc = L.svt_jpeg_xs_decoder_init(API_MAJOR, API_MINOR, ctypes.byref(self.api), buf.ctypes.data, len(buf),
                                        ctypes.byref(self.config))
        if rc:
            raise JxsError(rc, "decoder_init")

    def close(self):
        if self.api.private_ptr:
            lib().svt_jpeg_xs_decoder_close(ctypes.byref(self.api))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def image_frame(self, planes):
        """A Frame whose image buffer is the given output planes."""
        f = Frame()
        for c, p in enumerate(planes):
            f.image.data_yuv[c] = p.ctypes.data
            f.image.stride[c] = p.shape[1]
            f.image.alloc_size[c] = p.nbytes
        return f

    def new_planes(self):
        cfg = self.config
        dt = np.uint8 if cfg.bit_depth <= 8 else np.uint16
        return [np.zeros((cfg.components[c].height, cfg.components[c].width), dtype=dt)
                for c in range(cfg.components_num)]

    def send(self, buf, planes, ctx=None, blocking=1):
        """svt_jpeg_xs_decoder_send_frame of the codestream in numpy array buf
        into planes (both must stay alive until get returns the frame)."""
        f = self.image_frame(planes)
        f.bitstream.buffer = buf.ctypes.data
        f.bitstream.allocation_size = len(buf)
        f.bitstream.used_size = len(buf)
        f.user_prv_ctx_ptr = ctx
        return lib().svt_jpeg_xs_decoder_send_frame(ctypes.byref(self.api), ctypes.byref(f), blocking)

    def get(self, blocking=1):
        f = Frame()
        rc = lib().svt_jpeg_xs_decoder_get_frame(ctypes.byref(self.api), ctypes.byref(f), blocking)
        return rc, f

    def decode(self, codestream):
        cfg = self.config
        dt = np.uint8 if cfg.bit_depth <= 8 else np.uint16
        planes = [np.zeros((cfg.components[c].height, cfg.components[c].width), dtype=dt)
                  for c in range(cfg.components_num)]
        buf = np.frombuffer(codestream, dtype=np.uint8).copy()
        f = Frame()
        for c, p in enumerate(planes):
            f.image.data_yuv[c] = p.ctypes.data
            f.image.stride[c] = p.shape[1]
            f.image.alloc_size[c] = p.nbytes
        f.bitstream.buffer = buf.ctypes.data
        f.bitstream.allocation_size = len(buf)
        f.bitstream.used_size = len(buf)
        rc = lib().svt_jpeg_xs_decoder_send_frame(ctypes.byref(self.api), ctypes.byref(f), 1)
        if rc:
            return rc, []
        out = Frame()
        rc = lib().svt_jpeg_xs_decoder_get_frame(ctypes.byref(self.api), ctypes.byref(out), 1)
        return rc, planes


def _spec(spec, enc, planes, strides, out):
    spec.encoder = ctypes.pointer(enc.api)
    for c, (p, s) in enumerate(zip(planes, strides)):
        spec.d_planes[c] = p
        spec.strides[c] = s
    spec.d_out = out


def encode_noalloc(enc, planes, strides, out, d_status, stream=0):
    """svt_jpeg_xs_amd_encode_noalloc: one frame (device plane pointers,
    strides, device output pointer) on `stream`, its error flags written to
    the device u32 at d_status.  Returns the entry point's status."""
    spec = FrameSpec()
    _spec(spec, enc, planes, strides, out)
    return lib().svt_jpeg_xs_amd_encode_noalloc(ctypes.byref(spec), stream, d_status)


class Plan:
    """A device-resident batch (svt_jpeg_xs_amd_plan_*).  `frames` is a list
    of (encoder, [device plane pointers], [strides], device output pointer)."""

    def __init__(self, frames):
        self.frames = frames
        specs = self._specs_of(frames)
        self._specs = specs
        self.footprint = lib().svt_jpeg_xs_amd_plan_footprint(specs, len(frames))
        self.handle = lib().svt_jpeg_xs_amd_plan_create(specs, len(frames))
        if not self.handle:
            raise RuntimeError("svt_jpeg_xs_amd_plan_create failed")

    @staticmethod
    def _specs_of(frames):
        specs = (FrameSpec * len(frames))()
        for i, (enc, planes, strides, out) in enumerate(frames):
            _spec(specs[i], enc, planes, strides, out)
        return specs

    @staticmethod
    def footprint_of(frames):
        """Device bytes a plan (or a serving batch) of these frames takes."""
        return lib().svt_jpeg_xs_amd_plan_footprint(Plan._specs_of(frames), len(frames))

    def run(self, stream=0, use_graph=False):
        rc = lib().svt_jpeg_xs_amd_plan_run(self.handle, stream, 1 if use_graph else 0)
        if rc:
            raise JxsError(rc, "plan_run")

    def run_timed(self, stream=0):
        ms = (ctypes.c_float * STAGES)()
        rc = lib().svt_jpeg_xs_amd_plan_run_timed(self.handle, stream, ms)
        if rc:
            raise JxsError(rc, "plan_run_timed")
        return list(ms)

    def status(self):
        errs = (ctypes.c_uint32 * len(self.frames))()
        rc = lib().svt_jpeg_xs_amd_plan_status(self.handle, errs)
        return rc, list(errs)

    def precinct_offsets(self, frame):
        ptr = ctypes.c_void_p()
        stride = ctypes.c_uint32()
        rc = lib().svt_jpeg_xs_amd_plan_precinct_offsets(self.handle, frame, ctypes.byref(ptr), ctypes.byref(stride))
        if rc:
            raise JxsError(rc, "plan_precinct_offsets")
        return ptr.value, stride.value

    def destroy(self):
        if self.handle:
            lib().svt_jpeg_xs_amd_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Server:
    """The serving loop (svt_jpeg_xs_amd_server_*): frames of many streams,
    resident in HBM, admitted by HBM footprint and encoded by continuous
    batching on up to `max_inflight` HIP streams."""

    def __init__(self, hbm_budget, max_batch=64, max_inflight=2):
        cfg = ServerConfig(hbm_budget, max_batch, max_inflight)
        self.handle = lib().svt_jpeg_xs_amd_server_create(ctypes.byref(cfg))
        if not self.handle:
            raise RuntimeError("svt_jpeg_xs_amd_server_create failed")
        self._keep = {}

    @staticmethod
    def now_ns():
        return lib().svt_jpeg_xs_amd_server_now_ns()

    def submit(self, enc, planes, strides, out, tag):
        """Queues one frame: (encoder, [device plane pointers], [strides],
        device output pointer); returns the library's status code."""
        spec = FrameSpec()
        _spec(spec, enc, planes, strides, out)
        return lib().svt_jpeg_xs_amd_server_submit(self.handle, ctypes.byref(spec), tag)

    def poll(self, max_n=256, blocking=False):
        """Completed frames as dicts (tag, status, batch_frames, submit_ns,
        launch_ns, done_ns)."""
        buf = (Completion * max_n)()
        n = lib().svt_jpeg_xs_amd_server_poll(self.handle, buf, max_n, 1 if blocking else 0)
        return [{f: getattr(buf[i], f) for f, _ in Completion._fields_} for i in range(n)]

    def drain(self, count):
        """Blocks until `count` completions have been collected."""
        out = []
        while len(out) < count:
            got = self.poll(max(1, min(1024, count - len(out))), blocking=True)
            if not got:
                break
            out += got
        return out

    def stats(self):
        st = ServerStats()
        rc = lib().svt_jpeg_xs_amd_server_stats(self.handle, ctypes.byref(st))
        if rc:
            raise JxsError(rc, "server_stats")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def destroy(self):
        if self.handle:
            lib().svt_jpeg_xs_amd_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class DecodePlan:
    """A device-resident decode batch (svt_jpeg_xs_amd_decode_plan_*).
    `frames` is a list of (decoder, codestream device pointer, size,
    [plane device pointers], [strides], precinct table pointer or None,
    table stride)."""

    def __init__(self, frames):
        self.n = len(frames)
        specs = (DecodeSpec * self.n)()
        for i, (dec, cs, size, planes, strides, prec, pstride) in enumerate(frames):
            specs[i].decoder = ctypes.pointer(dec.api)
            specs[i].d_codestream = cs
            specs[i].codestream_size = size
            for c, (p, s) in enumerate(zip(planes, strides)):
                specs[i].d_planes[c] = p
                specs[i].strides[c] = s
            specs[i].d_precinct_offsets = prec
            specs[i].precinct_offsets_stride = pstride or 1
        self._specs = specs
        self._decoders = [f[0] for f in frames]
        self.handle = lib().svt_jpeg_xs_amd_decode_plan_create(specs, self.n)
        if not self.handle:
            raise RuntimeError("svt_jpeg_xs_amd_decode_plan_create failed")

    def run(self, stream=0):
        rc = lib().svt_jpeg_xs_amd_decode_plan_run(self.handle, stream)
        if rc:
            raise JxsError(rc, "decode_plan_run")

    def status(self):
        errs = (ctypes.c_uint32 * self.n)()
        rc = lib().svt_jpeg_xs_amd_decode_plan_status(self.handle, errs)
        return rc, list(errs)

    def destroy(self):
        if self.handle:
            lib().svt_jpeg_xs_amd_decode_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
