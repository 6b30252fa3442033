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
KERNELS = {"dwt1": "k_dwt", "dwt2": "k_dwt", "rate": "k_rate", "pack": "k_pack"}


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
            env = dict(os.environ, JXS_REF_SIMD="1")
            cmd = [O.REF_ENC, raw, os.path.join(td, "out.jxs"), str(W), str(H), str(DEPTH), str(FMT), str(BPP), "1",
                   str(DV), str(DH), str(len(sample)), str(cores)]
            t0 = time.time()
            r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=300)
            if r.returncode == 0:
                res = json.loads(r.stdout.strip().splitlines()[-1])
                return {"value": round(res["fps"], 3), "unit": "frames/s", "cores": cores, "kind": "reference",
                        "sample": "%d frames of the 3840x2160 4:2:2 10-bit bpp3 workload, reference "
                                  "SvtJpegxs (AVX2/AVX-512 kernels, lp=%d), %.1f s wall" %
                                  (len(sample), cores, time.time() - t0),
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
                   "frames_per_step": nf, "width": W, "height": H, "bit_depth": DEPTH, "bpp": BPP},
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
                     "algorithmic_bytes_per_launch": bytes_per_frame,
                     "pipeline_frac": round(bytes_per_frame * fps_plain / 1e9 / HBM_PEAK_GBS, 5)},
        "cpu_baseline": cpu,
        "api_host_buffers": api,
        "wall_s_plain": round(wall_plain, 3),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
