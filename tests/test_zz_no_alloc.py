"""svt_jpeg_xs_amd_encode_noalloc against the allocating plan path.

The entry point encodes one frame whose work block lives in the library's
fixed device array instead of a fresh allocation.  Two calls in a row, the
second with a smaller frame, must each give the codestream and error flags of
svt_jpeg_xs_amd_plan_create + _run + _status for the same frame (and the
oracle's codestream); a frame larger than the array is refused before any
work is enqueued.
"""
import numpy as np
import pytest

from jxs_util import load_oracle, load_package, load_synth

# 1920x1080 4:2:2 8-bit bpp 5 (32 MiB of work), then 640x360 4:4:4 10-bit bpp 3
CASES = [(1920, 1080, 8, 3, 5, 2, 31), (640, 360, 10, 4, 3, 1, 32)]


def _frame(torch, dev, cfg):
    S, O, P = load_synth(), load_oracle(), load_package()
    W, H, D, F, bpp, V, seed = cfg
    planes = S.make_frame(W, H, D, F, frame=seed)
    e, ref = O.encode(O.make_params(W, H, D, F, bpp, 1, V), planes)
    assert e == 0
    enc = P.Encoder(W, H, D, F, bpp, decomp_v=V, verbose=0)
    d_in = [torch.from_numpy(p.view(np.int16) if p.dtype == np.uint16 else p).to(dev) for p in planes]
    return enc, d_in, ref


@pytest.mark.gpu
def test_noalloc_equals_plan_path_twice():
    import torch
    P = load_package()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    status = torch.full((4,), 7, dtype=torch.int32, device=dev)
    for k, cfg in enumerate(CASES):
        print("noalloc call", k, cfg, flush=True)
        enc, d_in, ref = _frame(torch, dev, cfg)
        n = (enc.bytes_per_frame + 3) & ~3
        out_na = torch.full((n,), 0xA5, dtype=torch.uint8, device=dev)
        out_pl = torch.zeros(n, dtype=torch.uint8, device=dev)
        ptrs, strides = [t.data_ptr() for t in d_in], [t.shape[1] for t in d_in]
        rc = P.encode_noalloc(enc, ptrs, strides, out_na.data_ptr(), status.data_ptr(), stream.cuda_stream)
        assert rc == 0
        plan = P.Plan([(enc, ptrs, strides, out_pl.data_ptr())])
        plan.run(stream.cuda_stream)
        torch.cuda.synchronize()
        assert int(status[0]) == 0 and int(status[1]) == 7
        assert plan.status()[0] == 0
        got_na = out_na.cpu().numpy().tobytes()[:len(ref)]
        got_pl = out_pl.cpu().numpy().tobytes()[:len(ref)]
        assert got_na == got_pl, "call %d: differs from the plan path" % k
        assert got_na == ref, "call %d: differs from the oracle" % k
        plan.destroy()
        enc.close()


@pytest.mark.gpu
def test_noalloc_refuses_frames_beyond_its_array():
    import torch
    P = load_package()
    print("noalloc refusal", flush=True)
    dev = torch.device("cuda:0")
    enc = P.Encoder(3840, 2160, 10, 3, 3, verbose=0)
    plane = torch.zeros((2160, 3840), dtype=torch.int16, device=dev)
    chroma = torch.zeros((2160, 1920), dtype=torch.int16, device=dev)
    out = torch.zeros((enc.bytes_per_frame + 3) & ~3, dtype=torch.uint8, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    ptrs = [plane.data_ptr(), chroma.data_ptr(), chroma.data_ptr()]
    rc = P.encode_noalloc(enc, ptrs, [3840, 1920, 1920], out.data_ptr(), status.data_ptr(), 0)
    assert rc == P.ERR_BAD_PARAMETER
    torch.cuda.synchronize()
    assert int(out.sum()) == 0
    enc.close()
