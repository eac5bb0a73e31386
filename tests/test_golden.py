"""The oracle against the committed golden fixtures (tests/golden/, written by make_golden.py):
the same lattice (node keys, held flags, weights), joined_to_held bit for bit, and the
motions and bond forces to 1e-9 of the stored answers (the stored and the recomputed solves
both reach 1e-13; their difference is the solve's conditioning times that)."""
import glob
import json
import os

import numpy as np
import pytest

import oracle as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        header = json.load(f)
    return header, dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def test_fixtures_present():
    assert {"config0_16_seed0", "block_3", "single_cell", "stone_cantilever_2m", "unheld_piece"} <= set(NAMES)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    h, g = load(name)
    L = orc.Lattice(0, g["sites"], g["mats"], g["held"])
    assert L.n_nodes == h["nodes"] and L.n_bonds == h["bonds"]
    assert (L.around == g["around"]).all() and (L.held == g["node_held"]).all()
    assert np.array_equal(L.m, g["m"])
    joined = L.joined_to_held()
    assert (joined == g["joined"]).all()
    gone = (joined == 0).astype(np.uint8)
    x, _ = L.motions(g["loads"], gone=gone, tol=h["tol"], max_iter=20000)
    scale = np.abs(g["x"]).max()
    assert np.abs(x - g["x"]).max() <= 1e-9 * scale
    assert (x[gone.astype(bool)] == 0).all() and (x[L.held.astype(bool)] == 0).all()
    if "forces" in g:
        f = L.moved(x)
        assert np.abs(f - g["forces"]).max() <= 1e-9 * np.abs(g["forces"]).max()
