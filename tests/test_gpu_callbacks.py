"""Callback-driven callers and the frame-pool API on the GPU.

The IMTL ST 2110-22 plugin drives the encoder from two threads: a sender that
calls send_picture blocking, and a getter that polls get_packet non-blocking
and, on an empty queue, sleeps until callback_get_data_available wakes it
(imtl-plugin/src/st22_svt_jpeg_xs_plugin.c:57-149, callback set at :226-229).
Its decoder session waits for the first get callback before its first poll
(:414-435).  EncApp / DecApp take their buffers from frame pools and release
every output through svt_jpeg_xs_frame_pool_release (EncAppMain.c:113-166,
292-348; DecAppMain.c:121-160, 440-489).

Every wait here is bounded on the host, so a missing callback fails the test
instead of hanging it.
"""
import ctypes
import threading

import numpy as np
import pytest

from jxs_util import load_oracle, load_package, load_synth

WAIT_S = 10.0


def _frames(W, H, D, F, n, first=0):
    S = load_synth()
    return [S.make_frame(W, H, D, F, frame=first + k) for k in range(n)]


def _oracle_codestreams(W, H, D, F, bpp, frames, **kw):
    O = load_oracle()
    out = []
    for planes in frames:
        e, cs = O.encode(O.make_params(W, H, D, F, bpp, **kw), planes)
        assert e == 0
        out.append(cs)
    return out


class _Waker:
    """The plugin's condition variable, with a counter of callbacks so that a
    wake-up that lands before the wait is not lost."""

    def __init__(self):
        self.cond = threading.Condition()
        self.calls = 0

    def __call__(self):
        with self.cond:
            self.calls += 1
            self.cond.notify_all()

    def wait_beyond(self, n):
        with self.cond:
            return self.cond.wait_for(lambda: self.calls > n, timeout=WAIT_S)


def _encoder_session(slice_mode):
    P = load_package()
    W, H, D, F, bpp, n = 256, 64, 10, 3, 3, 32
    frames = _frames(W, H, D, F, n)
    refs = _oracle_codestreams(W, H, D, F, bpp, frames)
    waker, sends = _Waker(), _Waker()
    enc = P.Encoder(W, H, D, F, bpp, verbose=0, slice_packetization=slice_mode, callback_get=waker,
                    callback_send=sends)
    outs = [np.zeros(enc.bytes_per_frame, dtype=np.uint8) for _ in range(n)]
    errors = []

    def sender():
        for k in range(n):
            rc = enc.send(frames[k], outs[k], ctx=k + 1, blocking=1)
            if rc:
                errors.append("send_picture %d returned 0x%08x" % (k, rc & 0xFFFFFFFF))
                return

    packets = []

    def getter():
        taken = 0
        total = n * (1 + (H + 15) // 16 if slice_mode else 1)
        while taken < total:
            rc, f = enc.get(blocking=0)
            if rc == P.ERR_EMPTY_QUEUE:
                assert f.user_prv_ctx_ptr is None
                if not waker.wait_beyond(taken):
                    errors.append("no callback_get_data_available after %d packets" % taken)
                    return
                continue
            if rc:
                errors.append("get_packet returned 0x%08x" % (rc & 0xFFFFFFFF))
                return
            packets.append((f.user_prv_ctx_ptr, f.bitstream.buffer, f.bitstream.used_size,
                            f.bitstream.last_packet_in_frame))
            taken += 1

    ts = [threading.Thread(target=sender), threading.Thread(target=getter)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=60)
        assert not t.is_alive(), "session thread did not finish"
    enc.close()
    assert not errors, errors
    assert sends.calls >= 1  # at init, then whenever a frame slot frees
    return refs, outs, packets, waker.calls


@pytest.mark.gpu
def test_encoder_imtl_session_pattern():
    """32 frames through a blocking sender and a non-blocking getter woken
    only by callback_get_data_available: every frame comes out, in order,
    equal to the oracle's codestream, with one callback per frame."""
    print("encoder imtl pattern", flush=True)
    refs, outs, packets, calls = _encoder_session(0)
    assert [p[0] for p in packets] == list(range(1, 33))
    for k, (ctx, buf, used, last) in enumerate(packets):
        assert last == 1 and used == len(refs[k])
        assert outs[k][:used].tobytes() == refs[k], "frame %d" % k
    assert calls == 32


@pytest.mark.gpu
def test_encoder_imtl_session_slice_packets():
    """The same session in slice packetization mode: the frame header, then
    one packet per slice, each announced by its own callback
    (FinalStageProcess.c:173-175, 208-209); the packets of a frame concatenate
    to the oracle's codestream."""
    print("encoder imtl pattern, slice packets", flush=True)
    refs, outs, packets, calls = _encoder_session(1)
    per = 1 + 64 // 16
    assert len(packets) == 32 * per and calls == 32 * per
    for k in range(32):
        chunk = packets[k * per:(k + 1) * per]
        assert all(p[0] == k + 1 for p in chunk)
        assert [p[3] for p in chunk] == [0] * (per - 1) + [1]
        data = b"".join(ctypes.string_at(p[1], p[2]) for p in chunk)
        assert data == refs[k], "frame %d" % k


@pytest.mark.gpu
def test_decoder_imtl_session_pattern():
    """The IMTL decoder session: the getter first sleeps until the first
    callback_get_data_available, then polls non-blocking; 16 codestreams sent
    blocking come back in order, equal to the oracle's decode."""
    print("decoder imtl pattern", flush=True)
    P, O = load_package(), load_oracle()
    W, H, D, F, bpp, n = 256, 64, 10, 3, 3, 16
    css = _oracle_codestreams(W, H, D, F, bpp, _frames(W, H, D, F, n, first=40))
    bufs = [np.frombuffer(cs, dtype=np.uint8).copy() for cs in css]
    waker = _Waker()
    dec = P.Decoder(css[0], verbose=0, callback_get=waker)
    outs = [dec.new_planes() for _ in range(n)]
    errors, got = [], []

    def sender():
        for k in range(n):
            rc = dec.send(bufs[k], outs[k], ctx=k + 1, blocking=1)
            if rc:
                errors.append("send_frame %d returned 0x%08x" % (k, rc & 0xFFFFFFFF))
                return

    def getter():
        if not waker.wait_beyond(0):
            errors.append("no first callback")
            return
        taken = 0
        while taken < n:
            rc, f = dec.get(blocking=0)
            if rc == P.ERR_EMPTY_QUEUE:
                if not waker.wait_beyond(taken):
                    errors.append("no callback_get_data_available after %d frames" % taken)
                    return
                continue
            if rc:
                errors.append("get_frame returned 0x%08x" % (rc & 0xFFFFFFFF))
                return
            got.append(f.user_prv_ctx_ptr)
            taken += 1

    ts = [threading.Thread(target=getter), threading.Thread(target=sender)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=60)
        assert not t.is_alive(), "session thread did not finish"
    dec.close()
    assert not errors, errors
    assert got == list(range(1, n + 1))
    for k in range(n):
        e, ref = O.decode(css[k])
        assert e == 0
        for a, b in zip(outs[k], ref):
            assert np.array_equal(a, b), "frame %d" % k


def _copy_into(image, planes):
    for c, p in enumerate(planes):
        assert image.alloc_size[c] >= p.nbytes
        ctypes.memmove(image.data_yuv[c], p.ctypes.data, p.nbytes)


@pytest.mark.gpu
@pytest.mark.parametrize("slice_mode", [0, 1])
def test_encoder_through_frame_pool(slice_mode):
    """EncApp's buffer handling (EncAppMain.c:113-166, 292-348): input images
    and codestream buffers come from one frame pool of 3; the app releases
    after every packet, and a part goes back only when marked releasable
    (ImageBuffer.c:306-318).  No buffer is handed out while a frame still
    holds it, and every codestream equals the oracle's."""
    print("frame pool, slice mode", slice_mode, flush=True)
    P = load_package()
    L = P.lib()
    W, H, D, F, bpp, n = 256, 64, 10, 3, 3, 24
    frames = _frames(W, H, D, F, n, first=70)
    refs = _oracle_codestreams(W, H, D, F, bpp, frames)
    enc = P.Encoder(W, H, D, F, bpp, verbose=0, slice_packetization=slice_mode)
    pool = L.svt_jpeg_xs_frame_pool_alloc(ctypes.byref(enc.config), enc.bytes_per_frame, 3)
    assert pool
    assert not L.svt_jpeg_xs_frame_pool_alloc(ctypes.byref(enc.config), enc.bytes_per_frame, 0)
    lock = threading.Lock()
    held = set()
    errors = []

    def sender():
        for k in range(n):
            f = P.Frame()
            rc = L.svt_jpeg_xs_frame_pool_get(pool, ctypes.byref(f), 1)
            if rc:
                errors.append("pool_get %d" % rc)
                return
            with lock:
                for key in (("img", f.image.data_yuv[0]), ("bs", f.bitstream.buffer)):
                    if key in held:
                        errors.append("buffer %s handed out twice" % (key,))
                    held.add(key)
            _copy_into(f.image, frames[k])
            f.user_prv_ctx_ptr = k + 1
            rc = L.svt_jpeg_xs_encoder_send_picture(ctypes.byref(enc.api), ctypes.byref(f), 1)
            if rc:
                errors.append("send_picture %d" % rc)
                return

    t = threading.Thread(target=sender)
    t.start()
    got = {}
    done = 0
    while done < n and not errors:
        f = P.Frame()
        rc = L.svt_jpeg_xs_encoder_get_packet(ctypes.byref(enc.api), ctypes.byref(f), 1)
        assert rc == 0
        k = f.user_prv_ctx_ptr
        got.setdefault(k, b"")
        got[k] += ctypes.string_at(f.bitstream.buffer, f.bitstream.used_size)
        if f.bitstream.last_packet_in_frame:
            with lock:
                held.discard(("img", f.image.data_yuv[0]))
                # the codestream buffer starts at the frame's first packet
                held.discard(("bs", f.bitstream.buffer - (len(got[k]) - f.bitstream.used_size)))
            done += 1
        L.svt_jpeg_xs_frame_pool_release(pool, ctypes.byref(f))
        if f.bitstream.last_packet_in_frame:
            assert not f.image.release_ctx_ptr and not f.bitstream.release_ctx_ptr
    t.join(timeout=60)
    assert not t.is_alive()
    enc.close()
    L.svt_jpeg_xs_frame_pool_free(pool)
    assert not errors, errors
    assert sorted(got) == list(range(1, n + 1))
    for k in range(n):
        assert got[k + 1] == refs[k], "frame %d" % k


@pytest.mark.gpu
def test_frame_pool_release_rules():
    """svt_jpeg_xs_frame_pool_get / _release (ImageBuffer.c:248-318): a pool
    of one hands out nothing more until its parts are released; a frame not
    marked releasable returns nothing; a second release of a released frame
    returns nothing; image-only and bitstream-only pools work alone."""
    print("frame pool rules", flush=True)
    P = load_package()
    L = P.lib()
    cfg = P.ImageConfig()
    cfg.width, cfg.height, cfg.bit_depth, cfg.format, cfg.components_num = 64, 16, 10, P.FORMAT_YUV422, 3
    for c, (w, h) in enumerate([(64, 16), (32, 16), (32, 16)]):
        cfg.components[c].width, cfg.components[c].height, cfg.components[c].byte_size = w, h, w * h * 2
    pool = L.svt_jpeg_xs_frame_pool_alloc(ctypes.byref(cfg), 1000, 1)
    a, b = P.Frame(), P.Frame()
    assert L.svt_jpeg_xs_frame_pool_get(pool, ctypes.byref(a), 0) == 0
    assert a.image.stride[0] == 64 and a.image.stride[1] == 32 and a.bitstream.allocation_size == 1000
    assert L.svt_jpeg_xs_frame_pool_get(pool, ctypes.byref(b), 0) == P.ERR_EMPTY_QUEUE
    a.image.ready_to_release = 0
    a.bitstream.ready_to_release = 0
    L.svt_jpeg_xs_frame_pool_release(pool, ctypes.byref(a))
    assert L.svt_jpeg_xs_frame_pool_get(pool, ctypes.byref(b), 0) == P.ERR_EMPTY_QUEUE
    a.image.ready_to_release = 1
    L.svt_jpeg_xs_frame_pool_release(pool, ctypes.byref(a))  # image back, bitstream still held
    assert not a.image.release_ctx_ptr and a.bitstream.release_ctx_ptr
    assert L.svt_jpeg_xs_frame_pool_get(pool, ctypes.byref(b), 0) == P.ERR_EMPTY_QUEUE
    a.bitstream.ready_to_release = 1
    L.svt_jpeg_xs_frame_pool_release(pool, ctypes.byref(a))
    L.svt_jpeg_xs_frame_pool_release(pool, ctypes.byref(a))  # nothing left to return
    assert L.svt_jpeg_xs_frame_pool_get(pool, ctypes.byref(b), 0) == 0
    assert L.svt_jpeg_xs_frame_pool_get(pool, ctypes.byref(a), 0) == P.ERR_EMPTY_QUEUE
    L.svt_jpeg_xs_frame_pool_free(pool)
    img_only = L.svt_jpeg_xs_frame_pool_alloc(ctypes.byref(cfg), 0, 2)
    bs_only = L.svt_jpeg_xs_frame_pool_alloc(None, 500, 2)
    assert img_only and bs_only
    assert not L.svt_jpeg_xs_frame_pool_alloc(None, 0, 2)
    assert L.svt_jpeg_xs_frame_pool_get(img_only, ctypes.byref(a), 0) == 0 and a.image.data_yuv[0]
    assert L.svt_jpeg_xs_frame_pool_get(bs_only, ctypes.byref(b), 0) == 0 and b.bitstream.allocation_size == 500
    L.svt_jpeg_xs_frame_pool_free(img_only)
    L.svt_jpeg_xs_frame_pool_free(bs_only)
