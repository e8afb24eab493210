"""Writes tests/golden/bresenham_walks.json: voxel lists of hand-picked and seeded
segments, from a pure-Python transcription of the error recurrence of the
reference projector (3d/mat_mult.h:241-359: major axis = first of x, y, z with
the largest |d|; errors start at 2*m - L; a minor step when the error is > 0).
This generator shares no code with the C++ oracle, so the fixture pins the
oracle's walk independently.  Run: python tests/golden/make_golden.py"""
import json
import os

import numpy as np


def walk(p0, p1):
    x, y, z = p0
    dx, dy, dz = p1[0] - x, p1[1] - y, p1[2] - z
    xi, yi, zi = (-1 if dx < 0 else 1), (-1 if dy < 0 else 1), (-1 if dz < 0 else 1)
    mx, my, mz = abs(dx), abs(dy), abs(dz)
    out = [(x, y, z)]
    if mx >= my and mx >= mz:
        e1, e2 = 2 * my - mx, 2 * mz - mx
        for _ in range(mx):
            if e1 > 0:
                y += yi; e1 -= 2 * mx
            if e2 > 0:
                z += zi; e2 -= 2 * mx
            e1 += 2 * my; e2 += 2 * mz; x += xi
            out.append((x, y, z))
    elif my >= mx and my >= mz:
        e1, e2 = 2 * mx - my, 2 * mz - my
        for _ in range(my):
            if e1 > 0:
                x += xi; e1 -= 2 * my
            if e2 > 0:
                z += zi; e2 -= 2 * my
            e1 += 2 * mx; e2 += 2 * mz; y += yi
            out.append((x, y, z))
    else:
        e1, e2 = 2 * mx - mz, 2 * my - mz
        for _ in range(mz):
            if e1 > 0:
                x += xi; e1 -= 2 * mz
            if e2 > 0:
                y += yi; e2 -= 2 * mz
            e1 += 2 * mx; e2 += 2 * my; z += zi
            out.append((x, y, z))
    return out


def main():
    grid = (16, 12, 10)
    segs = [((0, 0, 0), (15, 11, 9)), ((15, 0, 0), (0, 11, 0)), ((3, 3, 3), (3, 3, 3)),
            ((2, 5, 1), (9, 5, 1)), ((2, 5, 1), (2, 11, 1)), ((2, 5, 1), (2, 5, 9)),
            ((0, 0, 0), (6, 6, 6)), ((5, 0, 9), (0, 5, 4)), ((1, 2, 3), (13, 7, 9)),
            ((14, 11, 0), (0, 4, 9)), ((0, 0, 0), (7, 7, 0)), ((0, 11, 0), (11, 0, 0))]
    rng = np.random.default_rng(2024)
    for _ in range(40):
        a = tuple(int(rng.integers(0, g)) for g in grid)
        b = tuple(int(rng.integers(0, g)) for g in grid)
        segs.append((a, b))
    recs = []
    for a, b in segs:
        vox = walk(a, b)
        lin = [x + y * grid[0] + z * grid[0] * grid[1] for x, y, z in vox]
        recs.append({"start": list(a), "end": list(b), "lin": lin})
    out = {"grid": list(grid), "segments": recs}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bresenham_walks.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(path, len(recs))


if __name__ == "__main__":
    main()
