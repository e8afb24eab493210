"""The kernels' closed-form Bresenham walk, emulated on the CPU with the same
float32 reciprocal and one-step fix-up, against the oracle's restatement of the
reference recurrence (3d/mat_mult.h:241-359).  Runs without a GPU."""
import numpy as np
import pytest

from tval3_amd import synth
import tv3_oracle as O


def closed_form_walk(x0, y0, z0, x1, y1, z1, rvx, rvy):
    d = np.array([x1 - x0, y1 - y0, z1 - z0])
    m = np.abs(d)
    strides = np.array([1, rvx, rvx * rvy])
    inc = np.where(d < 0, -1, 1) * strides
    if m[0] >= m[1] and m[0] >= m[2]:
        maj, a, b = 0, 1, 2
    elif m[1] >= m[0] and m[1] >= m[2]:
        maj, a, b = 1, 0, 2
    else:
        maj, a, b = 2, 0, 1
    L = int(m[maj])
    inv = np.float32(1.0) / np.float32(2 * L) if L > 0 else np.float32(0)
    j = np.arange(L + 1, dtype=np.int64)

    def k(mm):
        if mm == 0:
            return np.zeros_like(j)
        num = 2 * mm * j + L - 1
        den = 2 * L
        q = (num.astype(np.float32) * inv).astype(np.int64)   # float product, truncation
        q = np.where(q * den > num, q - 1, np.where((q + 1) * den <= num, q + 1, q))
        return q

    lin0 = x0 + y0 * rvx + z0 * rvx * rvy
    return lin0 + j * inc[maj] + k(int(m[a])) * inc[a] + k(int(m[b])) * inc[b]


@pytest.mark.parametrize("seed", range(4))
def test_closed_form_matches_recurrence(seed):
    rng = np.random.default_rng(seed)
    grid = [(64, 64, 64), (256, 256, 1), (7, 300, 20), (300, 5, 3)][seed]
    E, R = 40, 50
    pts = lambda k: [rng.integers(0, g, k) for g in grid]  # noqa: E731
    xe, ye, ze = pts(E)
    xr, yr, zr = pts(R)
    if seed == 1:  # planar rays plus axis-aligned and 45-degree ones
        xr[:5] = xe[:5]; yr[5:10] = ye[5:10]
        xr[10:15] = np.clip(xe[10:15] + 17, 0, 255); yr[10:15] = np.clip(ye[10:15] + 17, 0, 255)
    up = np.arange(1, E * R + 1, dtype=np.int32)
    geo = synth.GeometryArrays(1, up, xe, ye, ze, xr, yr, zr, grid, E * R)
    ptr, lin = O.Op.geometry(geo).walk()
    n = grid[0] * grid[1] * grid[2]
    assert lin.min() >= 0 and lin.max() < n
    for e in range(E):
        for r in range(R):
            row = e * R + r
            cf = closed_form_walk(xe[e], ye[e], ze[e], xr[r], yr[r], zr[r],
                                  grid[0], grid[1])
            np.testing.assert_array_equal(cf, lin[ptr[row]:ptr[row + 1]])
