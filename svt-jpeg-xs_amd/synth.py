"""Seeded synthetic test frames for the JPEG XS hot path.

Each plane is a sum of three seeded low-frequency 2D sinusoids (periods
64..512 px, amplitude (2^d-1)/8 around mid-grey), overwritten by 20 seeded
hard-edged rectangles, plus seeded Gaussian noise, rounded half away from
zero and clipped to [0, 2^d-1].  Seeds are splitmix64(base ^ frame<<8 ^ plane).

Arithmetic is float64 numpy on one machine image; tests pin the output by
SHA-256 so any drift between hosts is caught before parity is claimed.
"""
import hashlib

import numpy as np

BASE_SEED = 0x4A5853
MASK64 = (1 << 64) - 1

# ColourFormat_t values of the reference API (Source/API/SvtJpegxs.h:47-61).
FORMAT_YUV420 = 2
FORMAT_YUV422 = 3
FORMAT_YUV444 = 4
FORMAT_PACKED444 = 21


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def plane_seed(frame, plane, base=BASE_SEED):
    return splitmix64(base ^ (frame << 8) ^ plane)


def component_sizes(width, height, fmt):
    """(w, h) of each component, as svt_jpeg_xs_encoder_get_image_config."""
    if fmt == FORMAT_YUV422:
        return [(width, height), (width // 2, height), (width // 2, height)]
    if fmt == FORMAT_YUV420:
        return [(width, height), (width // 2, height // 2), (width // 2, height // 2)]
    if fmt in (FORMAT_YUV444, FORMAT_PACKED444):
        return [(width, height)] * 3
    raise ValueError("unsupported format %r" % fmt)


def make_plane(w, h, depth, seed, sigma=4.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    vmax = (1 << depth) - 1
    amp = vmax / 8.0
    mid = float(1 << (depth - 1))
    xs = np.arange(w, dtype=np.float64)
    ys = np.arange(h, dtype=np.float64)
    left = np.empty((h, 6), dtype=np.float64)
    right = np.empty((6, w), dtype=np.float64)
    for k in range(3):
        px = int(rng.integers(64, 513))
        py = int(rng.integers(64, 513))
        phi = float(rng.uniform(0.0, 2.0 * np.pi))
        ax = 2.0 * np.pi * xs / px
        by = 2.0 * np.pi * ys / py + phi
        # sin(ax + by) = sin(ax)cos(by) + cos(ax)sin(by): the three sinusoids
        # are one rank-6 product
        left[:, 2 * k] = np.cos(by)
        left[:, 2 * k + 1] = np.sin(by)
        right[2 * k] = np.sin(ax)
        right[2 * k + 1] = np.cos(ax)
    v = left @ right
    v *= amp
    v += mid
    for _ in range(20):
        rw = int(rng.integers(8, max(9, w // 4 + 1)))
        rh = int(rng.integers(8, max(9, h // 4 + 1)))
        x0 = int(rng.integers(0, max(1, w - rw + 1)))
        y0 = int(rng.integers(0, max(1, h - rh + 1)))
        val = float(rng.integers(0, vmax + 1))
        v[y0:y0 + rh, x0:x0 + rw] = val
    noise = rng.standard_normal((h, w))
    noise *= sigma
    v += noise
    # round half away from zero, then clip: every value below -0.5 clips to 0
    # either way, so floor(v + 0.5) gives the same result after the clip
    v += 0.5
    np.floor(v, out=v)
    np.clip(v, 0, vmax, out=v)
    return v.astype(np.uint8 if depth <= 8 else np.uint16)


def _frame_job(args):
    return make_frame(*args)


def _pool_map(mp, n, jobs):
    """pool.map with workers that end by themselves (close + join): a pool
    left by its context manager terminates its workers with SIGTERM, which a
    profiler's signal handler in the forked workers may not survive."""
    pool = mp.get_context("fork").Pool(n)
    try:
        return pool.map(_frame_job, jobs)
    finally:
        pool.close()
        pool.join()


def make_frames(width, height, depth, fmt, count, base=BASE_SEED, sigma=4.0, workers=None):
    """`count` frames with per-frame seeds, generated by a process pool."""
    import multiprocessing as mp
    import os
    jobs = [(width, height, depth, fmt, f, base, sigma) for f in range(count)]
    n = workers or min(16, os.cpu_count() or 1, count)
    if n <= 1:
        return [_frame_job(j) for j in jobs]
    return _pool_map(mp, n, jobs)


def make_frame(width, height, depth, fmt, frame=0, base=BASE_SEED, sigma=4.0):
    """Planes of one frame: a list of 2D arrays (planar formats) or a single
    (h, 3w) interleaved array for the packed format."""
    sizes = component_sizes(width, height, fmt)
    planes = [make_plane(w, h, depth, plane_seed(frame, c, base), sigma) for c, (w, h) in enumerate(sizes)]
    if fmt == FORMAT_PACKED444:
        return [np.ascontiguousarray(np.stack(planes, axis=-1).reshape(height, width * 3))]
    return planes


def flatten_rows(planes, height, period=16, flat=8):
    """Uneven content for rate-control tests: in every `period` rows of the
    picture the first `flat` are set to the plane's middle value (a plane of
    fewer rows is subsampled: its row y is picture row y * height / rows)."""
    out = []
    for p in planes:
        p = p.copy()
        sy = height // p.shape[0]
        rows = (np.arange(p.shape[0]) * sy) % period < flat
        p[rows] = (int(p.max()) + int(p.min())) // 2
        out.append(p)
    return out


def frame_bytes(planes):
    return b"".join(np.ascontiguousarray(p).tobytes() for p in planes)


def frame_sha256(planes):
    return hashlib.sha256(frame_bytes(planes)).hexdigest()


# The mixed workload: four stream kinds (width, height, depth, format, bpp),
# each coded with decomp_h 5 and decomp_v 1 or 2.
MIXED_KINDS = [
    (1280, 720, 8, FORMAT_YUV420, 1),    # 720p 4:2:0 8-bit
    (1920, 1080, 12, FORMAT_YUV444, 4),  # 1080p 4:4:4 12-bit
    (3840, 2160, 10, FORMAT_YUV444, 3),  # 2160p planar RGB 10-bit
    (1920, 540, 10, FORMAT_YUV422, 5),   # one interlaced 1080i field, 4:2:2 10-bit
]
MIXED_BASE = BASE_SEED + 3
SERVING_BASE = BASE_SEED + 4


def mixed_frames(count, base=MIXED_BASE):
    """(kind, decomp_v) of each frame of a mixed batch, drawn by seed."""
    rng = np.random.Generator(np.random.PCG64(splitmix64(base)))
    return [(int(rng.integers(0, len(MIXED_KINDS))), int(rng.integers(1, 3))) for _ in range(count)]


def serving_streams(count, base=SERVING_BASE):
    """(kind, decomp_v, fps) of each live stream: kind and levels from the
    mixed set, frame rate uniform in [30, 60]."""
    rng = np.random.Generator(np.random.PCG64(splitmix64(base)))
    return [(int(rng.integers(0, len(MIXED_KINDS))), int(rng.integers(1, 3)), float(rng.uniform(30.0, 60.0)))
            for _ in range(count)]


def poisson_arrivals(fps, duration_s, seed):
    """Arrival times (s) in [0, duration_s) of a Poisson process of rate fps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    t = float(rng.exponential(1.0 / fps))
    while t < duration_s:
        out.append(t)
        t += float(rng.exponential(1.0 / fps))
    return out


def make_frames_of(jobs, workers=None):
    """Frames for a list of (width, height, depth, fmt, frame, base) by a
    process pool."""
    import multiprocessing as mp
    import os
    jobs = [tuple(j) + (4.0,) for j in jobs]
    n = workers or min(16, os.cpu_count() or 1, len(jobs))
    if n <= 1:
        return [_frame_job(j) for j in jobs]
    return _pool_map(mp, n, jobs)
