"""Test scenes: the reference's experiment scenes (statics/examples/statics.rs),
restated as fill functions over cell centres, and small helpers.

Each scene is sampled exactly as `solids` does (statics/src/lattice.rs:87-102):
every site of the box lo..hi at the level's spacing takes what fills its
centre, x outermost, z innermost.
"""
import math

import numpy as np

from oracle import MASONRY, STEEL, STONE, WOOD, LEAVES, SOIL, UNIT  # noqa: F401


def inside(v, lo, hi):  # statics/examples/statics.rs:18-20
    return v >= lo - 1e-9 and v < hi - 1e-9


def solids(level, lo, hi, fill):
    """statics/src/lattice.rs:87-102: (sites, mats, held) of the filled sites."""
    s = UNIT * (1 << level)
    rng = [range(math.floor(lo[k] / s), math.ceil(hi[k] / s)) for k in range(3)]
    sites, mats, held = [], [], []
    for a in rng[0]:
        for b in rng[1]:
            for c in rng[2]:
                site = [a << level, b << level, c << level]
                centre = [(v + 0.5 * (1 << level)) * UNIT for v in site]
                f = fill(centre)
                if f is not None:
                    sites.append(site)
                    mats.append(f[0])
                    held.append(1 if f[1] else 0)
    return np.array(sites, np.int64).reshape(-1, 3), np.array(mats, np.float64).reshape(-1, 6), np.array(held, np.uint8)


def cantilever(mat, h, l):
    """A beam of section h x h, l long, out of a held wall (statics/examples/statics.rs:41-58)."""
    lo, hi = [-0.6, -0.6, -0.6], [l + 0.2, h + 0.6, h + 0.6]

    def fill(x):
        if inside(x[0], -0.5, 0.0) and inside(x[1], -0.5, h + 0.5) and inside(x[2], -0.5, h + 0.5):
            return STONE, True
        if inside(x[0], 0.0, l) and inside(x[1], 0.0, h) and inside(x[2], 0.0, h):
            return mat, False
        return None

    return lo, hi, fill


def portico(cut, slab):
    """A stone slab on four stone columns, `cut` of them cut through (statics/examples/statics.rs:87-111)."""
    w, high, c = 6.0, 3.0, 0.5
    columns = [[0.0, 0.0], [0.0, w - c], [w - c, 0.0], [w - c, w - c]]
    lo, hi = [-0.5, -0.5, -0.5], [w + 0.5, high + slab + 0.5, w + 0.5]

    def fill(x):
        if inside(x[1], -0.25, 0.0) and inside(x[0], -0.5, w + 0.5) and inside(x[2], -0.5, w + 0.5):
            return STONE, True
        if inside(x[1], high, high + slab) and inside(x[0], 0.0, w) and inside(x[2], 0.0, w):
            return STONE, False
        for k, col in enumerate(columns):
            if inside(x[1], 0.0, high) and inside(x[0], col[0], col[0] + c) and inside(x[2], col[1], col[1] + c):
                return (STONE, False) if (k < 4 - cut or not inside(x[1], 1.0, 1.0 + UNIT)) else None
        return None

    return lo, hi, fill


def tree(lean, side, depth):
    """A leaning tree cut `depth` of its width (statics/examples/statics.rs:60-85)."""
    r, h, crown = 0.3, 8.0, 2.0
    axis = [math.sin(math.radians(lean)), math.cos(math.radians(lean)), 0.0]
    middle = [a * (h + crown * 0.8) for a in axis]
    lo, hi = [-4.0, -0.5, -3.0], [6.0, h + 2.0 * crown + 0.5, 3.0]

    def fill(x):
        if x[1] < 0.0:
            return (STONE, True) if (abs(x[0]) < 1.0 and abs(x[2]) < 1.0 and x[1] >= -0.4) else None
        along = sum(x[k] * axis[k] for k in range(3))
        t = min(max(along, 0.0), h)
        off = [x[k] - axis[k] * t for k in range(3)]
        if math.sqrt(sum(o * o for o in off)) < r and along <= h:
            cut = inside(x[1], 0.5, 0.5 + UNIT) and side * off[0] > r - 2.0 * r * depth
            return None if cut else (WOOD, False)
        d = math.sqrt(sum((x[k] - middle[k]) ** 2 for k in range(3)))
        return (LEAVES, False) if d < crown else None

    return lo, hi, fill


def block(nx, ny, nz, held_layer=True, mat=STONE):
    """A box of full cells on a held layer (y = -1): sites, mats, held."""
    sites, held = [], []
    for x in range(nx):
        for y in range(-1 if held_layer else 0, ny):
            for z in range(nz):
                sites.append([x, y, z])
                held.append(1 if y < 0 else 0)
    sites = np.array(sites, np.int64)
    return sites, np.tile(np.array(mat, np.float64), (len(sites), 1)), np.array(held, np.uint8)
