"""Writes the golden fixtures of the solve path: small structures with the oracle's tightly
converged answers, stored as numpy archives (no pickles) with a JSON header each.

Each fixture holds its inputs (cells: sites, materials, held) and the oracle's outputs:
node keys (around), held flags and weights of the built lattice; joined_to_held; the
motions under the fixture's loads solved to 1e-13 (bodies::motions); and, for the small
ones, the bond forces (bodies::moved). The oracle is this repository's C++ restatement of statics/src/{model,
lattice, bodies}.rs; the reference itself is Rust and cannot run here. The fixtures pin the
oracle against regressions (tests/test_golden.py) and the device against fixed numbers
(tests/test_gpu_golden.py).

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "parallax-world_amd"), os.path.join(ROOT, "tests")]

import oracle as orc  # noqa: E402
import scenes  # noqa: E402
import statics_hip as sh  # noqa: E402
from harness import mats_from  # noqa: E402

TOL = 1e-13


def generated(n, seed):
    """Config 0's rule at n^3: pillars 2-6 wide at Poisson spacing 8, slabs every 8-16 with 20%
    2x2 holes, unit stiffness (e = 1, g = 0.4), ground layer held."""
    occ, _ = sh.generate(n, n, n, seed=seed, sigma=0.0, pillar_min=2, pillar_max=6)
    sites, held, e, g = sh.cells_from_occ(occ)
    return sites.astype(np.int64), mats_from(e, g), held


def block(n):
    """An n^3 block of unit stiffness on a held layer."""
    sites, held = [], []
    for x in range(n):
        for y in range(n):
            for z in range(n + 1):
                sites.append((x, y, z))
                held.append(1 if z == 0 else 0)
    sites = np.array(sites, np.int64)
    return sites, mats_from(np.ones(len(sites)), 0.4 * np.ones(len(sites))), np.array(held, np.uint8)


def unheld_piece():
    """A held 4 x 4 pad with a column on it, and a separate 2 x 2 x 2 piece floating beside it
    (not joined to anything held: gone)."""
    sites, held = [], []
    for x in range(4):
        for y in range(4):
            sites.append((x, y, 0))
            held.append(1)
    for z in range(1, 4):
        sites.append((1, 1, z))
        held.append(0)
    for x in range(6, 8):
        for y in range(6, 8):
            for z in range(2, 4):
                sites.append((x, y, z))
                held.append(0)
    sites = np.array(sites, np.int64)
    return sites, mats_from(np.ones(len(sites)), 0.4 * np.ones(len(sites))), np.array(held, np.uint8)


def cases():
    for seed in range(3):
        yield f"config0_16_seed{seed}", generated(16, seed), "gravity_z"
    yield "block_3", block(3), "gravity_z"
    yield "single_cell", block(1), "gravity_z"
    s, m, h = scenes.solids(0, *scenes.cantilever(orc.STONE, 0.5, 2.0))
    yield "stone_cantilever_2m", (s, m, h), "gravity_y"
    yield "unheld_piece", unheld_piece(), "gravity_z"


def loads_of(L, kind):
    f = np.zeros((L.n_nodes, 6))
    if kind == "gravity_z":  # weight 1 per free cell of density 1, a sixth per face, -z
        f[:, 2] = -L.m / orc.UNIT ** 3
    else:  # forces' loads (statics/src/bodies.rs:408)
        f = L.gravity()
    return f


def main():
    for name, (sites, mats, held), kind in cases():
        L = orc.Lattice(0, sites, mats, held)
        joined = L.joined_to_held()
        gone = (joined == 0).astype(np.uint8)
        f = loads_of(L, kind)
        x, st = L.motions(f, gone=gone, tol=TOL, max_iter=20000)
        arrays = dict(sites=sites, mats=mats, held=held, around=L.around, node_held=L.held, m=L.m, joined=joined,
                      loads=f, x=x)
        if L.n_bonds <= 20000:  # bond forces for the small cases only (they dominate the size)
            arrays["forces"] = L.moved(x)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
        header = {"name": name, "generator": "tests/golden/make_golden.py", "oracle": "oracle/statics_oracle.cpp",
                  "loads": kind, "tol": TOL, "iterations": st["iters"], "levels": st["levels"],
                  "cells": int(len(sites)), "nodes": int(L.n_nodes), "bonds": int(L.n_bonds),
                  "joined": int(joined.sum()), "max_abs_x": float(np.abs(x).max())}
        with open(os.path.join(HERE, f"{name}.json"), "w") as fh:
            json.dump(header, fh, indent=1)
        print(header)


if __name__ == "__main__":
    main()
