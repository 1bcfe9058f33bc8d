"""Inputs of the reference's own hot-path tests, rebuilt with the oracle's generator.

Sizes, times, velocities, NaN placement, far-out shifts and curved tables follow
tests/test_main.cpp:389-470 (test_backend_contract), :475-548 (test_shift_tables)
and :553-622 (test_cuda_algorithm). The random values come from the reference
tests' own xorshift64 + Box-Muller generator (test_main.cpp:28-36), restated in the
oracle (ora_test_normals).
"""
from __future__ import annotations

import math

import numpy as np

from oracle import pyoracle as po


def lround(x: float) -> int:
    r = math.trunc(x)
    if abs(x - r) >= 0.5:
        r += 1 if x > 0 else -1
    return int(r)


def contract_case(constant: bool, w=37, h=23, n=7, seed=4242, min_count=5):
    """test_backend_contract: symmetric times with x.5 shifts, 13x13 velocities of
    2 px/h plus (400, 0) far outside the frame."""
    t = np.array([(i - (n - 1) / 2.0) * 0.25 for i in range(n)])
    frames = np.ones((n, h, w), np.float32) if constant else po.test_normals(seed, (n, h, w))
    vel = np.array([(j * 2.0, k * 2.0) for j in range(-6, 7) for k in range(-6, 7)] + [(400.0, 0.0)])
    dx, dy = po.linear_shifts(t, vel)
    return dict(frames=frames, t=t, vel=vel, dx=dx, dy=dy, min_count=min_count)


def shift_table_case():
    """test_shift_tables: 41x29, 9 frames, 99 velocities, and the curved 150-motion
    table with a +500 jump at motion 77 and a +4 y-jump on every 17th motion in frame 3."""
    w, h, n = 41, 29, 9
    t = np.array([(i - 4) * 0.3 for i in range(n)])
    frames = po.test_normals(777, (n, h, w))
    vel = np.array([(j * 3.0, k * 2.5) for j in range(-5, 6) for k in range(-4, 5)])
    dx, dy = po.linear_shifts(t, vel)
    cdx = np.zeros((150, n), np.int32)
    cdy = np.zeros((150, n), np.int32)
    for j in range(150):
        for i in range(n):
            ti = t[i]
            ax, ay = (j % 7) - 3.0, (j % 5) - 2.0
            cdx[j, i] = lround((j % 13 - 6) * ti + ax * ti * ti) + (500 if j == 77 else 0)
            cdy[j, i] = lround((j % 11 - 5) * ti + ay * ti * ti) + (4 if (j % 17 == 0 and i == 3) else 0)
    return dict(frames=frames, t=t, vel=vel, dx=dx, dy=dy, cdx=cdx, cdy=cdy, min_count=6)


def cuda_algorithm_case(table: bool):
    """test_cuda_algorithm: NaN in frame 2 at flat index 100, 171 motions (not a
    multiple of any block size) with (400,0) and (0,-400); the curved table has a
    +500 x-jump at motion 77 and a -900 y-jump at motion 91."""
    w, h, n = 37, 23, 7
    t = np.array([(i - 3) * 0.25 for i in range(n)])
    frames = po.test_normals(99, (n, h, w))
    frames[2].reshape(-1)[100] = np.nan
    if not table:
        vel = np.array([(j * 2.0, k * 2.0) for j in range(-6, 7) for k in range(-6, 7)] + [(400.0, 0.0), (0.0, -400.0)])
        dx, dy = po.linear_shifts(t, vel)
    else:
        dx = np.zeros((150, n), np.int32)
        dy = np.zeros((150, n), np.int32)
        for j in range(150):
            for i in range(n):
                ti = t[i]
                dx[j, i] = lround((j % 13 - 6) * ti + ((j % 7) - 3.0) * ti * ti) + (500 if j == 77 else 0)
                dy[j, i] = lround((j % 11 - 5) * ti + ((j % 5) - 2.0) * ti * ti) - (900 if j == 91 else 0)
    return dict(frames=frames, t=t, dx=dx, dy=dy, min_count=5)


def all_stack_cases():
    """(name, frames, dx, dy, min_count) for every contract case above."""
    out = []
    for c in (False, True):
        k = contract_case(c)
        out.append((f"contract-{'constant' if c else 'random'}", k["frames"], k["dx"], k["dy"], k["min_count"]))
    s = shift_table_case()
    out.append(("shift-linear", s["frames"], s["dx"], s["dy"], s["min_count"]))
    out.append(("shift-curved", s["frames"], s["cdx"], s["cdy"], s["min_count"]))
    for tb in (False, True):
        k = cuda_algorithm_case(tb)
        out.append((f"cuda-{'table' if tb else 'linear'}", k["frames"], k["dx"], k["dy"], k["min_count"]))
    return out


def same_bits(a: np.ndarray, b: np.ndarray) -> bool:
    """Bit identity of float maps (NaN compares by bits too)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a.view(np.uint32), b.view(np.uint32))
