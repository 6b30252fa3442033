"""The first vertical 5/3 level as the transform kernel computes it, on
unscaled samples, against the oracle's lifting (Dwt.c restated) on the scaled
samples (NltEnc.c:33-71: x' = x * 2^s - 2^19, s = 20 - depth).  The kernel
forms D = 2 x(2i+1) - x(2i) - x(2i+2) on 16-bit sample pairs and relies on
  hf(i) = C * 2^(s-1),  lf(i) = (8 x(2i) + P + C2) * 2^(s-3) - 2^19,
with C = D(i) (0 without a high row), P = D(i-1) (C at the top row pair),
C2 = C (P without a high row), and x(2i+2) := x(2i) past the end.  This pins
that identity for every depth the API accepts, every column length up to 40
and the extreme sample values."""
import numpy as np
import pytest

from jxs_util import load_oracle


def raw_vertical(x, depth, mistake=None):
    s = 20 - depth
    n = len(x)
    lf = np.zeros((n + 1) // 2, dtype=np.int64)
    hf = np.zeros(n // 2, dtype=np.int64)
    for i in range((n + 1) // 2):
        prev, h, nxt = i > 0, i < n // 2, 2 * i + 2 < n
        c = int(x[2 * i])
        d = int(x[2 * i + 1]) if h else c
        if mistake == "clamped_end":
            e = int(x[min(2 * i + 2, n - 1)])
        else:
            e = int(x[2 * i + 2]) if nxt else c
        a = int(x[2 * i - 2]) if prev else c
        b = int(x[2 * i - 1]) if prev else c
        dc, dp = 2 * d - c - e, 2 * b - a - c
        assert -(1 << 15) <= dc < (1 << 15) and -(1 << 15) <= dp < (1 << 15)  # 16-bit lanes
        C = dc if h else 0
        P = dp if prev else (0 if mistake == "top_zero" else C)
        C2 = C if (h or mistake == "no_high_row") else P
        lf[i] = ((8 * c + P + C2) << (s - 3)) - (1 << 19)
        if h:
            hf[i] = C << (s - 1)
    return lf, hf


@pytest.mark.parametrize("depth", [8, 10, 12, 14])
def test_unscaled_vertical_step_equals_oracle_lifting(depth):
    O = load_oracle()
    rng = np.random.default_rng(1000 + depth)
    vmax = (1 << depth) - 1
    s = 20 - depth
    for n in range(1, 41):
        for trial in range(12):
            if trial == 0:
                x = np.full(n, vmax)
            elif trial == 1:
                x = np.zeros(n, dtype=np.int64)
            elif trial == 2:
                x = np.where(np.arange(n) % 2 == 0, 0, vmax)  # extreme high-pass
            elif trial == 3:
                x = np.where(np.arange(n) % 2 == 0, vmax, 0)
            else:
                x = rng.integers(0, vmax + 1, n)
            scaled = (x.astype(np.int64) << s) - (1 << 19)
            if n >= 2:
                lf_o, hf_o = O.fwd53(scaled)
            else:
                lf_o, hf_o = scaled.astype(np.int32), np.zeros(0, dtype=np.int32)
            lf_r, hf_r = raw_vertical(x, depth)
            assert np.array_equal(lf_o, lf_r), (depth, n, trial)
            assert np.array_equal(hf_o, hf_r), (depth, n, trial)


@pytest.mark.parametrize("mistake", ["clamped_end", "top_zero", "no_high_row"])
def test_planted_mistakes_are_caught(mistake):
    """Each edge rule of the identity matters: replacing the symmetric
    extension past the end by a clamp, the top row pair's P by 0, or the
    no-high-row case's C2 by C makes some column disagree with the oracle."""
    O = load_oracle()
    rng = np.random.default_rng(7)
    depth, s = 10, 10
    caught = False
    for n in range(2, 12):
        for _ in range(8):
            x = rng.integers(0, 1 << depth, n)
            lf_o, hf_o = O.fwd53((x.astype(np.int64) << s) - (1 << 19))
            lf_r, hf_r = raw_vertical(x, depth, mistake)
            caught |= not (np.array_equal(lf_o, lf_r) and np.array_equal(hf_o, hf_r))
    assert caught, mistake
