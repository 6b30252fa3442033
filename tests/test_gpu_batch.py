"""Batched encode: one launch sequence over heterogeneous frames (different
resolution, format, depth, levels, quantiser) through the device-resident
plan entry points, eager and replayed from a HIP graph, each frame's
codestream against the oracle."""
import numpy as np
import pytest

from jxs_util import load_oracle, load_package, load_synth

# (width, height, depth, format, bpp, decomp_v, decomp_h, quantization, significance[, rate control mode])
MIX = [
    (256, 64, 10, 3, 3, 2, 5, 0, 1),
    (200, 200, 8, 4, 3, 1, 5, 0, 1),
    (256, 128, 8, 2, 2, 2, 5, 0, 1),
    (96, 64, 14, 21, 2, 1, 5, 0, 1),
    (160, 90, 12, 4, 4, 2, 5, 1, 1),
    (130, 66, 9, 3, 4, 1, 2, 0, 1),
    (128, 64, 11, 4, 3, 0, 3, 0, 1),
    (160, 90, 10, 3, 5, 2, 5, 0, 0),
    (256, 64, 10, 3, 3, 2, 5, 0, 1, 0),
    (200, 200, 8, 4, 3, 1, 5, 0, 1, 0),
]


def _setup(torch, dev):
    S, O, P = load_synth(), load_oracle(), load_package()
    items = []
    for k, cfg in enumerate(MIX):
        W, H, D, F, bpp, V, Hh, q, sig = cfg[:9]
        rcm = cfg[9] if len(cfg) > 9 else 1
        planes = S.make_frame(W, H, D, F, frame=20 + k)
        e, ref = O.encode(O.make_params(W, H, D, F, bpp, 1, V, Hh, q, 16, sig, rcm), planes)
        assert e == 0
        enc = P.Encoder(W, H, D, F, bpp, 1, V, Hh, q, 16, sig, rate_control_mode=rcm)
        d_in = [torch.from_numpy(p.view(np.int16) if p.dtype == np.uint16 else p).to(dev) for p in planes]
        d_out = torch.zeros((enc.bytes_per_frame + 3) & ~3, dtype=torch.uint8, device=dev)
        items.append((enc, d_in, d_out, ref))
    specs = [(enc, [t.data_ptr() for t in d_in], [t.shape[1] for t in d_in], d_out.data_ptr())
             for enc, d_in, d_out, _ in items]
    return P, items, specs


def _check(items, tag):
    for k, (_, _, d_out, ref) in enumerate(items):
        got = d_out.cpu().numpy().tobytes()[:len(ref)]
        assert got == ref, "%s: frame %d %s differs" % (tag, k, MIX[k])


@pytest.mark.gpu
def test_heterogeneous_batch_eager_and_graph_replays():
    import torch
    dev = torch.device("cuda:0")
    print("batch of", len(MIX), "frames", flush=True)
    P, items, specs = _setup(torch, dev)
    plan = P.Plan(specs)
    stream = torch.cuda.current_stream(dev)
    plan.run(stream.cuda_stream)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    _check(items, "eager")
    for rep in range(3):
        print("graph replay", rep, flush=True)
        for _, _, d_out, _ in items:
            d_out.zero_()
        # eager work of the program between the replays
        scratch = torch.arange(1 << 16, device=dev, dtype=torch.int32).sum()
        torch.cuda.synchronize()
        assert int(scratch) > 0
        plan.run(stream.cuda_stream, use_graph=True)
        torch.cuda.synchronize()
        assert plan.status()[0] == 0
        _check(items, "replay %d" % rep)
    plan.destroy()


@pytest.mark.gpu
def test_batch_equals_single_frame_plans():
    import torch
    dev = torch.device("cuda:0")
    print("single-frame plans", flush=True)
    P, items, specs = _setup(torch, dev)
    stream = torch.cuda.current_stream(dev)
    plans = [P.Plan([s]) for s in specs]
    for p in plans:
        p.run(stream.cuda_stream)
    torch.cuda.synchronize()
    for p in plans:
        assert p.status()[0] == 0
        p.destroy()
    _check(items, "single")


@pytest.mark.gpu
def test_large_batch_shadow_words_match_oracle_and_single_frames():
    """A launch with more packing units than one 4K frame has (24 frames of
    320x180 4:2:2 10-bit, about 42,000 units) takes the packer's shadow-word
    path and its join kernel; every codestream must equal the oracle's and the
    one-frame plan's (which takes the direct path)."""
    import torch
    dev = torch.device("cuda:0")
    S, O, P = load_synth(), load_oracle(), load_package()
    W, H, D, F, bpp = 320, 180, 10, 3, 3
    n = 24
    print("batch of", n, "frames", W, H, flush=True)
    enc = P.Encoder(W, H, D, F, bpp, 1, 2, 5, verbose=0)
    prm = O.make_params(W, H, D, F, bpp, 1, 2, 5)
    lc = enc.bytes_per_frame
    refs, d_in = [], []
    for k in range(n):
        planes = S.make_frame(W, H, D, F, frame=500 + k)
        e, ref = O.encode(prm, planes)
        assert e == 0
        refs.append(ref)
        d_in.append([torch.from_numpy(p.view(np.int16)).to(dev) for p in planes])
    stride = (lc + 3) & ~3
    out_b = torch.zeros((n, stride), dtype=torch.uint8, device=dev)
    out_1 = torch.zeros((n, stride), dtype=torch.uint8, device=dev)
    spec = lambda k, out: (enc, [t.data_ptr() for t in d_in[k]], [t.shape[1] for t in d_in[k]], out[k].data_ptr())
    batch = P.Plan([spec(k, out_b) for k in range(n)])
    stream = torch.cuda.current_stream(dev)
    for rep in range(2):
        out_b.zero_()
        batch.run(stream.cuda_stream, use_graph=rep == 1)
        torch.cuda.synchronize()
        assert batch.status()[0] == 0
        host = out_b.cpu().numpy()
        for k in range(n):
            assert host[k, :lc].tobytes() == refs[k], "batch frame %d (replay %d) differs" % (k, rep)
    for k in (0, n - 1):
        plan = P.Plan([spec(k, out_1)])
        plan.run(stream.cuda_stream)
        torch.cuda.synchronize()
        assert plan.status()[0] == 0
        assert out_1[k, :lc].cpu().numpy().tobytes() == refs[k]
        plan.destroy()
    batch.destroy()
    enc.close()


@pytest.mark.gpu
def test_autotuned_tiles_keep_codestreams():
    """The autotuner picks a transform width per job shape and the rate table
    kernel's waves per precinct (checking that every candidate gives the same
    codestream on its own test frame); encoding afterwards, plain and
    batched, still equals the oracle."""
    import torch
    dev = torch.device("cuda:0")
    S, O, P = load_synth(), load_oracle(), load_package()
    # shapes with one, two and five horizontal levels, 8- and 16-bit input
    cfgs = [(256, 64, 10, 3, 3, 2, 5), (200, 200, 8, 4, 3, 1, 5), (128, 64, 11, 4, 3, 0, 3)]
    items = []
    for k, (W, H, D, F, bpp, V, Hh) in enumerate(cfgs):
        enc = P.Encoder(W, H, D, F, bpp, 1, V, Hh, verbose=0)
        tiles = enc.autotune()
        print("autotune", (W, H, D, F, V, Hh), tiles, flush=True)
        assert tiles and all(t["frame_us"] > 0 for t in tiles)
        assert all(t["chunk"] in ((8, 16, 32) if t["stage"] == 0 else (1, 2, 4)) for t in tiles)
        assert [t["stage"] for t in tiles].count(1) == 1  # every config here is rate control mode 1
        planes = S.make_frame(W, H, D, F, frame=700 + k)
        e, ref = O.encode(O.make_params(W, H, D, F, bpp, 1, V, Hh), planes)
        assert e == 0
        rc, cs = enc.encode(planes)
        assert rc == 0 and cs == ref, "plain path after autotune, config %d" % k
        d_in = [torch.from_numpy(p.view(np.int16) if p.dtype == np.uint16 else p).to(dev) for p in planes]
        d_out = torch.zeros((enc.bytes_per_frame + 3) & ~3, dtype=torch.uint8, device=dev)
        items.append((enc, d_in, d_out, ref))
    plan = P.Plan([(enc, [t.data_ptr() for t in d_in], [t.shape[1] for t in d_in], d_out.data_ptr())
                   for enc, d_in, d_out, _ in items])
    plan.run(torch.cuda.current_stream(dev).cuda_stream)
    torch.cuda.synchronize()
    assert plan.status()[0] == 0
    for k, (enc, _, d_out, ref) in enumerate(items):
        assert d_out.cpu().numpy().tobytes()[:len(ref)] == ref, "batch after autotune, config %d" % k
    plan.destroy()
    for enc, _, _, _ in items:
        enc.close()
