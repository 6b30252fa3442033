"""Serving loop (svt_jpeg_xs_amd_server_*): frames of several live streams of
different resolution, format, depth and levels, submitted interleaved and
encoded by continuous batching on several HIP streams.  Each frame's
codestream must be byte-identical in the mix, alone (a one-frame plan) and to
the oracle; admission must keep the HBM in use within the budget."""
import numpy as np
import pytest

from jxs_util import load_oracle, load_package, load_synth

# (width, height, depth, format, bpp, decomp_v, decomp_h[, rate control mode]):
# one live stream each, the four kinds of the mixed serving workload at reduced
# size, and one stream with rate control mode 0
STREAMS = [
    (256, 144, 8, 2, 1, 2, 5),   # 4:2:0 8-bit bpp 1
    (192, 108, 12, 4, 4, 2, 5),  # 4:4:4 12-bit bpp 4
    (320, 180, 10, 4, 3, 1, 5),  # planar RGB 10-bit bpp 3
    (192, 54, 10, 3, 5, 2, 5),   # 4:2:2 10-bit bpp 5 (a field)
    (192, 108, 10, 3, 3, 2, 5, 0),  # 4:2:2 10-bit bpp 3, budget per precinct
]
FRAMES_PER_STREAM = 3

OK = 0
ERR_BAD_PARAMETER = -0x7FFFEFFB
ERR_INSUFFICIENT_RESOURCES = -0x7FFFF000


def _setup(torch, dev):
    S, O, P = load_synth(), load_oracle(), load_package()
    streams = []
    for k, cfg in enumerate(STREAMS):
        W, H, D, F, bpp, V, Hh = cfg[:7]
        rcm = cfg[7] if len(cfg) > 7 else 1
        enc = P.Encoder(W, H, D, F, bpp, 1, V, Hh, verbose=0, rate_control_mode=rcm)
        prm = O.make_params(W, H, D, F, bpp, 1, V, Hh, rc_mode=rcm)
        frames = []
        for f in range(FRAMES_PER_STREAM):
            planes = S.make_frame(W, H, D, F, frame=100 * k + f)
            e, ref = O.encode(prm, planes)
            assert e == 0
            d_in = [torch.from_numpy(p.view(np.int16) if p.dtype == np.uint16 else p).to(dev) for p in planes]
            d_out = torch.zeros((enc.bytes_per_frame + 3) & ~3, dtype=torch.uint8, device=dev)
            frames.append((d_in, d_out, ref))
        streams.append((enc, frames))
    return P, streams


def _spec(enc, d_in, d_out):
    return enc, [t.data_ptr() for t in d_in], [t.shape[1] for t in d_in], d_out.data_ptr()


def _interleaved(streams):
    """(tag, stream, frame) in arrival order: round robin over the streams."""
    order = []
    for f in range(FRAMES_PER_STREAM):
        for s in range(len(streams)):
            order.append((s * 1000 + f, s, f))
    return order


def _serve(P, streams, budget, max_batch, max_inflight):
    srv = P.Server(budget, max_batch, max_inflight)
    order = _interleaved(streams)
    for tag, s, f in order:
        enc, frames = streams[s]
        d_in, d_out, _ = frames[f]
        d_out.zero_()
    import torch
    torch.cuda.synchronize()
    for tag, s, f in order:
        enc, frames = streams[s]
        d_in, d_out, _ = frames[f]
        rc = srv.submit(*_spec(enc, d_in, d_out), tag)
        assert rc == OK, (tag, rc)
    done = srv.drain(len(order))
    stats = srv.stats()
    srv.destroy()
    return order, done, stats


def _check(streams, order, done, what):
    assert len(done) == len(order), what
    assert sorted(c["tag"] for c in done) == sorted(t for t, _, _ in order)
    for c in done:
        assert c["status"] == OK, (what, c)
        assert c["submit_ns"] <= c["launch_ns"] <= c["done_ns"]
    for tag, s, f in order:
        d_in, d_out, ref = streams[s][1][f]
        got = d_out.cpu().numpy().tobytes()[:len(ref)]
        assert got == ref, "%s: stream %d %s frame %d differs" % (what, s, STREAMS[s], f)


@pytest.mark.gpu
def test_mix_equals_alone_and_oracle():
    import torch
    dev = torch.device("cuda:0")
    print("server: %d streams x %d frames" % (len(STREAMS), FRAMES_PER_STREAM), flush=True)
    P, streams = _setup(torch, dev)
    # alone: every frame in a one-frame plan of its own
    stream = torch.cuda.current_stream(dev)
    for enc, frames in streams:
        for d_in, d_out, ref in frames:
            plan = P.Plan([_spec(enc, d_in, d_out)])
            plan.run(stream.cuda_stream)
            torch.cuda.synchronize()
            assert plan.status()[0] == 0
            plan.destroy()
            assert d_out.cpu().numpy().tobytes()[:len(ref)] == ref
    # in the mix: interleaved submissions, batches of up to 5 on 2 streams
    print("server: mixed batches", flush=True)
    order, done, stats = _serve(P, streams, 256 << 20, 5, 2)
    _check(streams, order, done, "mix")
    assert stats["completed"] == len(order)
    assert stats["hbm_in_use"] == 0
    assert stats["hbm_peak"] <= stats["hbm_budget"]
    for enc, _ in streams:
        enc.close()


@pytest.mark.gpu
def test_admission_by_hbm_budget():
    """A budget that holds about one frame's work at a time: every frame is
    still encoded, exactly, and the memory in use never exceeds the budget."""
    import torch
    dev = torch.device("cuda:0")
    print("server: tight budget", flush=True)
    P, streams = _setup(torch, dev)
    fp = max(P.Plan.footprint_of([_spec(enc, fr[0][0], fr[0][1])]) for enc, fr in streams)
    budget = fp + (fp // 2)
    order, done, stats = _serve(P, streams, budget, 8, 3)
    _check(streams, order, done, "tight budget")
    assert stats["hbm_peak"] <= budget
    # frames smaller than half the largest may share a launch; the rest go alone
    assert stats["batches"] >= len(order) // 3
    # a frame larger than the whole budget is refused at submission
    srv = P.Server(fp // 4, 4, 1)
    enc, frames = streams[np.argmax([P.Plan.footprint_of([_spec(e, f[0][0], f[0][1])]) for e, f in streams])]
    assert srv.submit(*_spec(enc, frames[0][0], frames[0][1]), 1) == ERR_INSUFFICIENT_RESOURCES
    srv.destroy()
    for enc, _ in streams:
        enc.close()


@pytest.mark.gpu
def test_continuous_batching_groups_waiting_frames():
    """Frames submitted while a batch runs are carried together by the next
    launches: fewer launches than frames, and the results stay exact."""
    import torch
    dev = torch.device("cuda:0")
    print("server: burst", flush=True)
    P, streams = _setup(torch, dev)
    order, done, stats = _serve(P, streams, 512 << 20, 16, 1)
    _check(streams, order, done, "burst")
    assert stats["batches"] < len(order)
    assert stats["largest_batch"] > 1
    assert max(c["batch_frames"] for c in done) == stats["largest_batch"]
    for enc, _ in streams:
        enc.close()


@pytest.mark.gpu
def test_submit_rejects_invalid_specs():
    import torch
    dev = torch.device("cuda:0")
    P = load_package()
    enc = P.Encoder(64, 32, 8, 4, 3, verbose=0)
    d_in = [torch.zeros((32, 64), dtype=torch.uint8, device=dev) for _ in range(3)]
    d_out = torch.zeros(enc.bytes_per_frame + 8, dtype=torch.uint8, device=dev)
    srv = P.Server(64 << 20, 4, 1)
    # codestream buffer not 4-byte aligned
    assert srv.submit(enc, [t.data_ptr() for t in d_in], [64] * 3, d_out.data_ptr() + 1, 7) == ERR_BAD_PARAMETER
    # missing plane
    assert srv.submit(enc, [d_in[0].data_ptr(), 0, d_in[2].data_ptr()], [64] * 3, d_out.data_ptr(), 8) == \
        ERR_BAD_PARAMETER
    # stride below the component width
    assert srv.submit(enc, [t.data_ptr() for t in d_in], [32] * 3, d_out.data_ptr(), 9) == ERR_BAD_PARAMETER
    assert srv.poll() == []
    assert srv.stats()["submitted"] == 0
    srv.destroy()
    enc.close()
