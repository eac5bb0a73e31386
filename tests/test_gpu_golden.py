"""The device against the committed golden fixtures (tests/golden/): joined_to_held bit for bit,
the solve (FP64 and mixed, tol 1e-10) within 1e-8 of the stored motions, and the bond forces
of the stored motions within 1e-12 of the stored ones."""
import glob
import json
import os

import numpy as np
import pytest

import statics_hip as sh
from harness import FACES

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


@pytest.mark.parametrize("name", NAMES)
def test_device_matches_fixture(ctx, name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        h = json.load(f)
    g = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    sites, mats, held = g["sites"], g["mats"], g["held"]
    ctx.upload(sites.astype(np.int32), held, mats[:, 1], mats[:, 2])
    around = g["around"]
    o = np.argmax(around[:, 1] != around[:, 0], axis=1)
    idx = ctx.lookup_faces(np.concatenate([o[:, None], around[:, 1]], axis=1).astype(np.int32))
    assert (idx >= 0).all() and len(idx) == ctx.nodes == h["nodes"]
    assert (ctx.node_held()[idx] == g["node_held"]).all()
    assert (ctx.joined_to_held()[idx] == g["joined"]).all()
    fd = np.zeros((ctx.nodes, 6))
    fd[idx] = g["loads"]
    scale = np.abs(g["x"]).max()
    for prec in (sh.FP64, sh.MIXED):
        x, st = ctx.solve(fd, precision=prec, tol=1e-10)
        assert st.status == 0
        assert np.abs(x[idx] - g["x"]).max() <= 1e-8 * scale, (prec, st.iters)
    if "forces" in g:
        xd = np.zeros((ctx.nodes, 6))
        xd[idx] = g["x"]
        f = ctx.bond_forces(xd)  # [cell][q][k][3], upload order
        # each stored bond: its beam's cell and pair from its nodes' sites (build's order)
        site_of = {tuple(s): i for i, s in enumerate(sites.tolist())}
        nb = len(g["forces"]) // 4
        # the oracle emits each free cell's beams in pair order, cell by cell in input order
        cells = [i for i in range(len(sites)) if not held[i]]
        k = 0
        got = np.zeros_like(g["forces"])
        for c in cells:
            for q, (a, b) in enumerate([(a, b) for a in range(6) for b in range(a + 1, 6)]):
                fa = tuple(sites[c] + FACES[a])
                fb = tuple(sites[c] + FACES[b])
                ha = fa in site_of and held[site_of[fa]]
                hb = fb in site_of and held[site_of[fb]]
                if ha and hb:
                    continue  # held-held beams are not built (statics/src/lattice.rs:154)
                got[4 * k:4 * k + 4] = f[c, q]
                k += 1
        assert k == nb
        assert np.abs(got - g["forces"]).max() <= 1e-12 * np.abs(g["forces"]).max()
