"""Benchmark of the MI355X JPEG XS encoder on one GPU.

Workload (one "step"): 64 seeded synthetic frames of 3840x2160 YUV 4:2:2
10-bit planar (u16), --bpp 3 --decomp_v 2 --decomp_h 5, every other
parameter at the reference's defaults, encoded one frame per launch sequence
(the plain GPU path) with the frames already resident in HBM.  `value` is
encoded frames/s of that plain path.  The same line reports the batched path
(all 64 frames in one launch sequence, replayed from a HIP graph), per-frame
latency percentiles, the roofline of the dominant kernel, bit-exactness
against the oracle, and the reference CPU encoder timed on this host.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tests"))

W, H, DEPTH, FMT, BPP, DV, DH = 3840, 2160, 10, 3, 3, 2, 5
NFRAMES = 64
HBM_PEAK_GBS = 8000.0
STAGES = ["dwt1", "dwt2", "rate", "pack"]
KERNELS = {"dwt1": "k_dwt", "dwt2": "k_dwt", "rate": "k_rate_tab + k_rate_plan", "pack": "k_pack"}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "encode_3840x2160_422_10b_bpp3_traffic.json")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic(stage):
    """HBM bytes per launch of one stage's kernel on this workload, from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised in profiles/ (each
    scaled by the counter's reading for a copy of known size), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        k = t["stages"][stage]
        return int(k["fetch_bytes"] + k["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def pct(xs, p):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(round((len(xs) - 1) * p / 100.0)))]


def cpu_baseline(frames, budget_s=20.0):
    """The reference encoder (built from source into oracle/_ref) on a bounded
    sample of the same frames, with as many threads as this host's share
    (at most 16); falls back to the single-threaded oracle port."""
    import numpy as np
    from jxs_util import load_oracle
    O = load_oracle()
    cores = min(16, os.cpu_count() or 1)
    if O.have_reference():
        sample = frames[:16]
        with tempfile.TemporaryDirectory() as td:
            raw = os.path.join(td, "in.raw")
            with open(raw, "wb") as f:
                for fr in sample:
                    for p in fr:
                        f.write(np.ascontiguousarray(p).tobytes())
            cmd = [O.REF_ENC, raw, os.path.join(td, "out.jxs"), str(W), str(H), str(DEPTH), str(FMT), str(BPP), "1",
                   str(DV), str(DH), str(len(sample)), str(cores)]
            # a first pass over the sample sizes the timed run: the sample's frames
            # sent in turn for about budget_s / 2 seconds (at most 400 passes)
            r = subprocess.run(cmd, capture_output=True, text=True, env=dict(os.environ, JXS_REF_SIMD="1"), timeout=300)
            if r.returncode == 0:
                fps0 = json.loads(r.stdout.strip().splitlines()[-1])["fps"]
                rep = max(1, min(400, int(fps0 * budget_s / 2 / len(sample))))
                env = dict(os.environ, JXS_REF_SIMD="1", JXS_REF_REPEAT=str(rep))
                t0 = time.time()
                r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=300)
            if r.returncode == 0:
                res = json.loads(r.stdout.strip().splitlines()[-1])
                return {"value": round(res["fps"], 3), "unit": "frames/s", "cores": cores, "kind": "reference",
                        "sample": "%d distinct frames of the 3840x2160 4:2:2 10-bit bpp3 workload sent %d times "
                                  "(%d frames), reference SvtJpegxs built from source (AVX2/AVX-512 C kernels; its "
                                  "three NASM routines run as their C twins, no assembler in this image), lp=%d "
                                  "threads on this host, %.1f s wall" %
                                  (len(sample), rep, res["frames"], cores, time.time() - t0),
                        "p50_ms": res["p50_ms"], "p99_ms": res["p99_ms"]}
            log("reference baseline failed rc=%d: %s" % (r.returncode, r.stderr[-300:]))
    p = O.make_params(W, H, DEPTH, FMT, BPP, 1, DV, DH)
    t0 = time.time()
    n = 0
    while n < len(frames) and time.time() - t0 < budget_s:
        O.encode(p, frames[n])
        n += 1
    dt = time.time() - t0
    return {"value": round(n / dt, 3), "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": "%d frames of the workload through the scalar C oracle, %.1f s" % (n, dt)}


def api_host_path(P, torch, frames, n):
    """Encodes n frames through svt_jpeg_xs_encoder_send_picture /
    get_packet with planes and codestream buffers in pinned host memory: a
    sender thread feeds the encoder (blocking send), this thread collects
    packets in order.  Includes the PCIe copies both ways."""
    import threading
    import numpy as np
    enc = P.Encoder(W, H, DEPTH, FMT, BPP, 1, DV, DH, verbose=0)
    nb = len(frames)
    pinned = [[torch.from_numpy(p.view(np.int16)).pin_memory() for p in fr] for fr in frames[:min(nb, 8)]]
    host = [[t.numpy().view(np.uint16) for t in fr] for fr in pinned]
    outs = [torch.empty(enc.bytes_per_frame, dtype=torch.uint8).pin_memory() for _ in range(4)]
    bufs = [o.numpy() for o in outs]
    errors = []

    def sender():
        for i in range(n):
            rc = enc.send(host[i % len(host)], bufs[i % len(bufs)])
            if rc:
                errors.append(rc)
                return

    # warm the slots once
    enc.send(host[0], bufs[0])
    enc.get()
    t0 = time.perf_counter()
    th = threading.Thread(target=sender)
    th.start()
    done = 0
    while done < n and not errors:
        rc, _ = enc.get()
        if rc:
            errors.append(rc)
            break
        done += 1
    th.join()
    dt = time.perf_counter() - t0
    enc.close()
    if errors:
        return {"error": int(errors[0])}
    return {"value": round(done / dt, 2), "unit": "frames/s", "frames": done,
            "note": "C API send_picture/get_packet, planes and codestream in pinned host memory, "
                    "PCIe copies included; up to 3 frames in flight"}


def _dev_planes(torch, dev, planes):
    import numpy as np
    return [torch.from_numpy(p.view(np.int16) if p.dtype == np.uint16 else p).to(dev) for p in planes]


def _spec(enc, d_in, out_ptr):
    return (enc, [t.data_ptr() for t in d_in], [t.shape[1] for t in d_in], out_ptr)


def _tiles(tiles):
    """Autotuner choices: 'width/levels(v) -> chunk' per transform job shape."""
    return ["rate/%d->%dw" % (t["width"], t["chunk"]) if t.get("stage") == 1 else
            "%d/%d%s->%d" % (t["width"], t["levels"], "v" if t["vertical"] else "", t["chunk"]) for t in tiles]


def _time_plans(torch, stream, plans, batch, steps):
    """Milliseconds of `steps` passes of the one-frame plans and of the
    batch plan (graph replay), HIP events on the launch stream."""
    sh = stream.cuda_stream
    for p in plans:
        p.run(sh)
    batch.run(sh, use_graph=True)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        for p in plans:
            p.run(sh)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_plain = ev0.elapsed_time(ev1)
    ev0.record(stream)
    for _ in range(steps):
        batch.run(sh, use_graph=True)
    ev1.record(stream)
    torch.cuda.synchronize()
    return ms_plain, ev0.elapsed_time(ev1)


def plain_vs_batched_1080p(P, S, torch, dev, steps, check, nframes=64):
    """1920x1080 4:2:2 8-bit, bpp 5, decomp_v 2, decomp_h 5: one frame per
    launch sequence against all frames in one launch sequence."""
    import numpy as np
    W1, H1, D1, F1, B1 = 1920, 1080, 8, 3, 5
    frames = S.make_frames(W1, H1, D1, F1, nframes)
    enc = P.Encoder(W1, H1, D1, F1, B1, 1, 2, 5, verbose=0)
    tiles = enc.autotune(torch.cuda.current_stream(dev).cuda_stream)
    ostride = (enc.bytes_per_frame + 3) & ~3
    d_in = [_dev_planes(torch, dev, fr) for fr in frames]
    out_p = torch.zeros((nframes, ostride), dtype=torch.uint8, device=dev)
    out_b = torch.zeros((nframes, ostride), dtype=torch.uint8, device=dev)
    plans = [P.Plan([_spec(enc, d_in[i], out_p[i].data_ptr())]) for i in range(nframes)]
    batch = P.Plan([_spec(enc, d_in[i], out_b[i].data_ptr()) for i in range(nframes)])
    stream = torch.cuda.current_stream(dev)
    ms_plain, ms_batch = _time_plans(torch, stream, plans, batch, steps)
    hp = out_p[:, :enc.bytes_per_frame].cpu().numpy()
    hb = out_b[:, :enc.bytes_per_frame].cpu().numpy()
    exact = bool(np.array_equal(hp, hb))
    if check:
        from jxs_util import load_oracle
        O = load_oracle()
        prm = O.make_params(W1, H1, D1, F1, B1, 1, 2, 5)
        for k in (0, nframes - 1):
            e, ref = O.encode(prm, frames[k])
            exact = exact and e == 0 and ref == hp[k].tobytes()
    fp = nframes * steps / (ms_plain / 1e3)
    fb = nframes * steps / (ms_batch / 1e3)
    for p in plans:
        p.destroy()
    batch.destroy()
    enc.close()
    return {"workload": "1920x1080 YUV422 8-bit planar, bpp 5, decomp_v 2, decomp_h 5, %d frames" % nframes,
            "plain_fps": round(fp, 1), "batched_fps": round(fb, 1), "batched_over_plain": round(fb / fp, 3),
            "tiles": _tiles(tiles), "bit_exact": bool(exact)}


def mixed_batch(P, S, torch, dev, steps, check, nframes=48):
    """The uneven batch: frames of four kinds (720p 4:2:0 8-bit bpp 1, 1080p
    4:4:4 12-bit bpp 4, 2160p planar RGB 10-bit bpp 3, 1920x540 4:2:2 10-bit
    bpp 5), decomp_v 1 or 2, drawn by seed, in one launch sequence."""
    import numpy as np
    draw = S.mixed_frames(nframes)
    jobs = []
    for i, (k, v) in enumerate(draw):
        W1, H1, D1, F1, _ = S.MIXED_KINDS[k]
        jobs.append((W1, H1, D1, F1, i, S.MIXED_BASE))
    frames = S.make_frames_of(jobs)
    encs, tiles = {}, []
    for k, v in sorted(set(draw)):
        W1, H1, D1, F1, B1 = S.MIXED_KINDS[k]
        encs[(k, v)] = P.Encoder(W1, H1, D1, F1, B1, 1, v, 5, verbose=0)
        tiles += encs[(k, v)].autotune(torch.cuda.current_stream(dev).cuda_stream)
    d_in = [_dev_planes(torch, dev, fr) for fr in frames]
    lcods = [encs[kv].bytes_per_frame for kv in draw]
    ostride = (max(lcods) + 3) & ~3
    out_p = torch.zeros((nframes, ostride), dtype=torch.uint8, device=dev)
    out_b = torch.zeros((nframes, ostride), dtype=torch.uint8, device=dev)
    plans = [P.Plan([_spec(encs[draw[i]], d_in[i], out_p[i].data_ptr())]) for i in range(nframes)]
    batch = P.Plan([_spec(encs[draw[i]], d_in[i], out_b[i].data_ptr()) for i in range(nframes)])
    stream = torch.cuda.current_stream(dev)
    ms_plain, ms_batch = _time_plans(torch, stream, plans, batch, steps)
    hp = out_p.cpu().numpy()
    hb = out_b.cpu().numpy()
    exact = all(np.array_equal(hp[i, :lcods[i]], hb[i, :lcods[i]]) for i in range(nframes))
    exact = exact and batch.status()[0] == 0
    if check:
        from jxs_util import load_oracle
        O = load_oracle()
        seen = set()
        for i, (k, v) in enumerate(draw):
            if k in seen:
                continue
            seen.add(k)
            W1, H1, D1, F1, B1 = S.MIXED_KINDS[k]
            e, ref = O.encode(O.make_params(W1, H1, D1, F1, B1, 1, v, 5), frames[i])
            exact = exact and e == 0 and ref == hb[i, :lcods[i]].tobytes()
    pixels = sum(S.MIXED_KINDS[k][0] * S.MIXED_KINDS[k][1] for k, _ in draw)
    in_bytes = sum(sum(p.nbytes for p in fr) for fr in frames) + sum(lcods)
    fp = nframes * steps / (ms_plain / 1e3)
    fb = nframes * steps / (ms_batch / 1e3)
    for p in plans:
        p.destroy()
    batch.destroy()
    for e in encs.values():
        e.close()
    counts = {}
    for k, v in draw:
        key = "%dx%d_%s_v%d" % (S.MIXED_KINDS[k][0], S.MIXED_KINDS[k][1], ["", "", "420", "422", "444"][
            S.MIXED_KINDS[k][3]], v)
        counts[key] = counts.get(key, 0) + 1
    return {"workload": "%d frames drawn by seed from 4 kinds (720p 420 8b bpp1, 1080p 444 12b bpp4, 2160p "
                        "planar RGB 10b bpp3, 1920x540 422 10b bpp5), decomp_v 1-2, decomp_h 5" % nframes,
            "mix": counts, "tiles": _tiles(tiles), "batched_fps": round(fb, 1), "plain_fps": round(fp, 1),
            "batched_gpixel_per_s": round(pixels * steps / (ms_batch / 1e3) / 1e9, 3),
            "batched_hbm_frac": round(in_bytes * steps / (ms_batch / 1e3) / 1e9 / HBM_PEAK_GBS, 5),
            "ms_per_batch": round(ms_batch / steps, 3), "bit_exact": bool(exact)}


def roundtrip_8k(P, S, torch, dev, steps, nframes=16, distinct=4):
    """7680x4320 packed RGB 14-bit, bpp 2, decomp_v 1, decomp_h 5 (noise sigma
    16): the frames encoded in one batch, then decoded in one batch (the
    decoder takes the encoder's precinct tables).  Frame 0's codestream and
    decoded planes are hashed against the reference encoder's and decoder's
    (tests/golden); the other frames repeat `distinct` generated inputs."""
    import hashlib
    import numpy as np
    W5, H5, D5, F5, B5 = 7680, 4320, 14, 21, 2
    frames = S.make_frames(W5, H5, D5, F5, distinct, sigma=16.0)
    enc = P.Encoder(W5, H5, D5, F5, B5, 1, 1, 5, verbose=0)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream
    tiles = enc.autotune(sh)
    lc = enc.bytes_per_frame
    d_in = [torch.from_numpy(fr[0].view(np.int16)).to(dev) for fr in frames]
    del frames
    d_cs = torch.zeros((nframes, (lc + 3) & ~3), dtype=torch.uint8, device=dev)
    plan = P.Plan([(enc, [d_in[i % distinct].data_ptr()], [3 * W5], d_cs[i].data_ptr()) for i in range(nframes)])
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    plan.run(sh, use_graph=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(steps):
        plan.run(sh, use_graph=True)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_enc = ev0.elapsed_time(ev1)
    enc_ok = plan.status()[0] == 0
    cs0 = d_cs[0, :lc].cpu().numpy().tobytes()
    dec = P.Decoder(cs0, verbose=0)
    d_out = [[torch.zeros((H5, W5), dtype=torch.int16, device=dev) for _ in range(3)] for _ in range(nframes)]
    specs = []
    for i in range(nframes):
        prec, pstride = plan.precinct_offsets(i)
        specs.append((dec, d_cs[i].data_ptr(), lc, [t.data_ptr() for t in d_out[i]], [W5] * 3, prec, pstride))
    dplan = P.DecodePlan(specs)
    dplan.run(sh)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(steps):
        dplan.run(sh)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_dec = ev0.elapsed_time(ev1)
    dec_ok = dplan.status()[0] == 0
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))["full"]["c5_7680x4320_rgbpacked_14b_bpp2_v1"]
    cs_ref = hashlib.sha256(cs0).hexdigest() == g["codestream_sha256"]
    dec0 = b"".join(t.cpu().numpy().tobytes() for t in d_out[0])
    dec_ref = hashlib.sha256(dec0).hexdigest() == g["decoded_sha256"]
    per_frame = 3 * W5 * H5 * 2 + lc
    fe = nframes * steps / (ms_enc / 1e3)
    fd = nframes * steps / (ms_dec / 1e3)
    dplan.destroy()
    plan.destroy()
    dec.close()
    enc.close()
    return {"workload": "7680x4320 packed RGB 14-bit, bpp 2, decomp_v 1, decomp_h 5, %d frames per batch, encode "
                        "then decode, inputs resident in HBM" % nframes,
            "tiles": _tiles(tiles), "encode_fps": round(fe, 1), "decode_fps": round(fd, 1),
            "encode_gpixel_per_s": round(fe * W5 * H5 / 1e9, 2), "decode_gpixel_per_s": round(fd * W5 * H5 / 1e9, 2),
            "encode_hbm_frac": round(fe * per_frame / 1e9 / HBM_PEAK_GBS, 4),
            "decode_hbm_frac": round(fd * per_frame / 1e9 / HBM_PEAK_GBS, 4),
            "frame_errors": int(not enc_ok) + int(not dec_ok),
            "codestream_equals_reference": bool(cs_ref), "decoded_equals_reference": bool(dec_ref)}


def serving(P, S, torch, dev, check, nstreams=32, seconds=3.0, budget_gib=16, max_batch=32, inflight=3):
    """The serving loop: nstreams live streams (kind and levels drawn from
    the mixed set, Poisson arrivals at 30-60 frames/s each) submitted in real
    time to one server; latency is submission to observed completion.  Then
    a burst of 8 frames per stream at once measures the served capacity."""
    import numpy as np
    POOL, RING = 2, 8
    streams = S.serving_streams(nstreams)
    kinds = sorted({k for k, _, _ in streams})
    jobs = [(S.MIXED_KINDS[k][0], S.MIXED_KINDS[k][1], S.MIXED_KINDS[k][2], S.MIXED_KINDS[k][3], 1000 * k + f,
             S.SERVING_BASE) for k in kinds for f in range(POOL)]
    frames = S.make_frames_of(jobs)
    pool = {}
    for (W1, H1, D1, F1, fr, _), planes in zip(jobs, frames):
        k = fr // 1000
        pool.setdefault(k, []).append(_dev_planes(torch, dev, planes))
    encs, outs, tuned = [], [], set()
    for k, v, _ in streams:
        W1, H1, D1, F1, B1 = S.MIXED_KINDS[k]
        e = P.Encoder(W1, H1, D1, F1, B1, 1, v, 5, verbose=0)
        if (k, v) not in tuned:
            e.autotune(torch.cuda.current_stream(dev).cuda_stream)
            tuned.add((k, v))
        encs.append(e)
        outs.append(torch.zeros((RING, (e.bytes_per_frame + 3) & ~3), dtype=torch.uint8, device=dev))
    # each (stream kind, levels, pool frame) encoded alone, in a one-frame plan
    expected = {}
    stream = torch.cuda.current_stream(dev)
    for s, (k, v, _) in enumerate(streams):
        if (k, v) in expected:
            continue
        got = []
        for f in range(POOL):
            scratch = torch.zeros_like(outs[s][0])
            plan = P.Plan([_spec(encs[s], pool[k][f], scratch.data_ptr())])
            plan.run(stream.cuda_stream)
            torch.cuda.synchronize()
            got.append(scratch[:encs[s].bytes_per_frame].cpu().numpy().tobytes())
            plan.destroy()
        expected[(k, v)] = got
    srv = P.Server(budget_gib << 30, max_batch, inflight)

    def submit(s, n, tag):
        k = streams[s][0]
        rc = srv.submit(encs[s], [t.data_ptr() for t in pool[k][n % POOL]], [t.shape[1] for t in pool[k][n % POOL]],
                        outs[s][n % RING].data_ptr(), tag)
        if rc:
            raise RuntimeError("server_submit rc=%d" % rc)

    # warm-up: one frame per stream
    for s in range(nstreams):
        submit(s, 0, s)
    srv.drain(nstreams)
    # real-time Poisson arrivals
    events = sorted((t, s) for s, (_, _, fps) in enumerate(streams)
                    for t in S.poisson_arrivals(fps, seconds, S.splitmix64(S.SERVING_BASE ^ (s + 1))))
    count = [0] * nstreams
    comps = []
    t0 = time.monotonic_ns() + 20_000_000
    for i, (t, s) in enumerate(events):
        target = t0 + int(t * 1e9)
        while True:
            now = time.monotonic_ns()
            if now >= target:
                break
            if target - now > 2_000_000:
                time.sleep((target - now) / 1e9 - 0.001)
        submit(s, count[s], i)
        count[s] += 1
        if (i & 15) == 15:
            comps += srv.poll(1024)
    comps += srv.drain(len(events) - len(comps))
    st_rt = srv.stats()
    lat = np.array([(c["done_ns"] - c["submit_ns"]) / 1e6 for c in comps])
    wait = np.array([(c["launch_ns"] - c["submit_ns"]) / 1e6 for c in comps])
    bsz = np.array([c["batch_frames"] for c in comps])
    errs = sum(1 for c in comps if c["status"] != 0)
    span = (max(c["done_ns"] for c in comps) - min(c["submit_ns"] for c in comps)) / 1e9
    exact = errs == 0 and len(comps) == len(events)

    def ring_ok():
        ok = True
        for s, (k, v, _) in enumerate(streams):
            host = outs[s].cpu().numpy()
            lc = encs[s].bytes_per_frame
            for j in range(min(RING, count[s])):
                ok = ok and host[j, :lc].tobytes() == expected[(k, v)][j % POOL]
        return ok

    exact = exact and ring_ok()
    # burst: RING frames per stream at once (every frame its own buffer)
    for o in outs:
        o.zero_()
    torch.cuda.synchronize()
    b0 = srv.now_ns()
    for n in range(RING):
        for s in range(nstreams):
            submit(s, n, (1 << 32) + n * nstreams + s)
    bc = srv.drain(RING * nstreams)
    b1 = max(c["done_ns"] for c in bc)
    st = srv.stats()
    count = [RING] * nstreams
    burst_exact = len(bc) == RING * nstreams and all(c["status"] == 0 for c in bc) and ring_ok()
    burst_lat = np.array([(c["done_ns"] - c["submit_ns"]) / 1e6 for c in bc])
    srv.destroy()
    for e in encs:
        e.close()
    mix = {}
    for k, v, _ in streams:
        key = "%dx%d_%s_v%d" % (S.MIXED_KINDS[k][0], S.MIXED_KINDS[k][1], ["", "", "420", "422", "444"][
            S.MIXED_KINDS[k][3]], v)
        mix[key] = mix.get(key, 0) + 1
    return {"workload": "%d live streams (kind and decomp_v drawn from the mixed set), Poisson arrivals at "
                        "30-60 frames/s each, %.1f s real time, frames resident in HBM" % (nstreams, seconds),
            "streams": mix, "offered_fps": round(sum(f for _, _, f in streams), 1), "frames": len(comps),
            "aggregate_fps": round(len(comps) / span, 1),
            "p50_latency_ms": round(float(np.percentile(lat, 50)), 4),
            "p99_latency_ms": round(float(np.percentile(lat, 99)), 4),
            "p999_latency_ms": round(float(np.percentile(lat, 99.9)), 4),
            "max_latency_ms": round(float(lat.max()), 4),
            "p99_queue_wait_ms": round(float(np.percentile(wait, 99)), 4),
            "mean_batch_frames": round(float(bsz.mean()), 2), "batches": int(st_rt["batches"]) - nstreams,
            "frame_errors": errs, "bit_exact_vs_alone": bool(exact),
            "hbm_budget_bytes": int(st["hbm_budget"]), "hbm_peak_bytes": int(st["hbm_peak"]),
            "max_batch": max_batch, "max_inflight": inflight,
            "loop": "host-side: one worker thread waits at most 20 us for work, reaps finished launch slots and "
                    "launches each batch's kernels on one of %d HIP streams; nothing runs persistently on the GPU"
                    % inflight,
            "burst": {"frames": len(bc), "served_fps": round(len(bc) / ((b1 - b0) / 1e9), 1),
                      "p99_latency_ms": round(float(np.percentile(burst_lat, 99)), 4),
                      "largest_batch": int(st["largest_batch"]), "bit_exact_vs_alone": bool(burst_exact)},
            "target_p99_ms": round(1000.0 / 60.0, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=NFRAMES)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--api-frames", type=int, default=32,
                    help="frames pushed through the C API with host buffers (PCIe-inclusive rate; 0: skip)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop run")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed batch and the 1080p comparison")
    ap.add_argument("--no-roundtrip", action="store_true", help="skip the 8K encode + decode round trip")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--streams", type=int, default=32)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB first: the reference point for the "
                         "FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every box here has one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from jxs_util import load_package, load_synth
    P = load_package()
    S = load_synth()
    dev = torch.device("cuda:0")
    nf = args.frames

    log("generating %d frames" % nf)
    frames = S.make_frames(W, H, DEPTH, FMT, nf)
    enc = P.Encoder(W, H, DEPTH, FMT, BPP, 1, DV, DH)
    log("autotuning transform tiles")
    tiles4k = enc.autotune(torch.cuda.current_stream(dev).cuda_stream)
    lcod = enc.bytes_per_frame
    in_bytes = sum(p.nbytes for p in frames[0])
    bytes_per_frame = in_bytes + lcod

    log("uploading frames to HBM")
    d_in = [[torch.from_numpy(p.view(np.int16) if p.dtype == np.uint16 else p).to(dev) for p in fr] for fr in frames]
    ostride = (lcod + 3) & ~3  # codestream buffers are 4-byte aligned
    d_out_plain = torch.zeros((nf, ostride), dtype=torch.uint8, device=dev)
    d_out_batch = torch.zeros((nf, ostride), dtype=torch.uint8, device=dev)

    def spec(i, out):
        return (enc, [t.data_ptr() for t in d_in[i]], [t.shape[1] for t in d_in[i]], out[i].data_ptr())

    plans = [P.Plan([spec(i, d_out_plain)]) for i in range(nf)]
    batch = P.Plan([spec(i, d_out_batch) for i in range(nf)])
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream

    if args.copy_calibration:
        n = args.copy_calibration << 20
        log("calibration copy of %d MiB" % args.copy_calibration)
        a = torch.ones(n // 4, dtype=torch.int32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    log("warmup")
    for _ in range(args.warmup):
        for p in plans:
            p.run(sh)
        batch.run(sh, use_graph=True)
    torch.cuda.synchronize()

    # plain path: one launch sequence per frame
    log("timing plain path")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.steps):
        for p in plans:
            p.run(sh)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall_plain = time.perf_counter() - t0
    ms_plain = ev0.elapsed_time(ev1)
    fps_plain = nf * args.steps / (ms_plain / 1e3)

    # batched path: all frames in one launch sequence (HIP graph replay)
    log("timing batched path")
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        batch.run(sh, use_graph=True)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_batch = ev0.elapsed_time(ev1)
    fps_batch = nf * args.steps / (ms_batch / 1e3)

    # per-frame latency of the plain path, one frame in flight at a time
    lat = []
    for p in plans:
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        p.run(sh)
        b.record(stream)
        b.synchronize()
        lat.append(a.elapsed_time(b))

    # per-kernel durations (HIP events on the launch stream) -> roofline
    stage_ms = np.zeros(len(STAGES))
    nt = min(nf, 16)
    for p in plans[:nt]:
        stage_ms += np.array(p.run_timed(sh))
    stage_ms /= nt
    dom = int(np.argmax(stage_ms))
    dom_ms = float(stage_ms[dom])
    achieved = bytes_per_frame / (dom_ms / 1e3) / 1e9

    errs = 0
    for p in plans:
        rc, _ = p.status()
        errs += rc != 0
    rc, _ = batch.status()
    errs += rc != 0
    plain_host = d_out_plain[:, :lcod].cpu().numpy()
    batch_host = d_out_batch[:, :lcod].cpu().numpy()
    same = bool(np.array_equal(plain_host, batch_host))
    exact = None
    if not args.no_check:
        from jxs_util import load_oracle
        O = load_oracle()
        prm = O.make_params(W, H, DEPTH, FMT, BPP, 1, DV, DH)
        exact = same and errs == 0
        for k in sorted({0, nf // 2, nf - 1}):
            e, ref = O.encode(prm, frames[k])
            exact = exact and e == 0 and ref == plain_host[k].tobytes()
        exact = bool(exact)

    # the C API with host buffers (send_picture / get_packet from pinned
    # memory, a sender thread and a receiver as EncApp does): PCIe-inclusive
    api = None
    if args.api_frames:
        api = api_host_path(P, torch, frames, args.api_frames)

    mixed = hd = srv = None
    if not args.no_mixed:
        log("1080p plain vs batched")
        hd = plain_vs_batched_1080p(P, S, torch, dev, args.steps, not args.no_check)
        log("mixed batch")
        mixed = mixed_batch(P, S, torch, dev, args.steps, not args.no_check)
    rt = None
    if not args.no_roundtrip:
        log("8K round trip")
        rt = roundtrip_8k(P, S, torch, dev, args.steps)
    if not args.no_serving:
        log("serving loop")
        srv = serving(P, S, torch, dev, not args.no_check, nstreams=args.streams, seconds=args.serve_seconds)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(frames)

    out = {
        "metric": "encoded frames/s + p99 frame latency at 1 GPU, batched vs plain, bit-exact",
        "value": round(fps_plain, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_plain / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic (seeded sinusoids + rectangles + Gaussian noise, per-frame seeds)",
        "config": {"workload": "3840x2160 YUV422 10-bit planar, bpp 3, decomp_v 2, decomp_h 5, %d frames, "
                               "one frame per launch, inputs resident in HBM" % nf,
                   "frames_per_step": nf, "width": W, "height": H, "bit_depth": DEPTH, "bpp": BPP,
                   "transform_tiles": _tiles(tiles4k)},
        "gpixel_per_s": round(fps_plain * W * H / 1e9, 3),
        "p50_frame_latency_ms": round(pct(lat, 50), 4),
        "p99_frame_latency_ms": round(pct(lat, 99), 4),
        "batched": {"value": round(fps_batch, 2), "unit": "frames/s", "frames_per_launch": nf,
                    "ms_per_step": round(ms_batch / args.steps, 3), "speedup_vs_plain": round(fps_batch / fps_plain, 3),
                    "hip_graph": True},
        "bit_exact": exact,
        "plain_equals_batched": same,
        "frame_errors": errs,
        "stage_ms_per_frame": {s: round(float(v), 5) for s, v in zip(STAGES, stage_ms)},
        "roofline": {"bound": "hbm", "kernel": KERNELS[STAGES[dom]], "stage": STAGES[dom],
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(STAGES[dom]),
                     "duration_us": round(dom_ms * 1e3, 3),
                     "duration_source": "HIP events around the kernel on its launch stream, mean of %d frames; "
                                        "frac from the kernel-trace mean instead is algorithmic_bytes_per_launch / "
                                        "(trace mean) / peak" % nt,
                     "algorithmic_bytes_per_launch": bytes_per_frame,
                     "pipeline_frac": round(bytes_per_frame * fps_plain / 1e9 / HBM_PEAK_GBS, 5)},
        "cpu_baseline": cpu,
        "api_host_buffers": api,
        "plain_vs_batched_1080p": hd,
        "mixed_batch": mixed,
        "roundtrip_8k": rt,
        "serving": srv,
        "wall_s_plain": round(wall_plain, 3),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
