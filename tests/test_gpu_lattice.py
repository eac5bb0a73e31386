"""Lattices as the reference holds them, through st_lattice_cells and st_upload, on the device:
a built lattice solves as the oracle's motions does, and a lattice made stronger beam by beam
by bodies::as_generated (statics/src/bodies.rs:604-625) and then edited, as the reference's
`edits` experiment does (statics/examples/statics.rs:326-336), settles round for round as the
oracle's settle, with the beams' own strengths."""
import numpy as np
import pytest

import oracle as orc
import statics_hip as sh
from harness import mats_from

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def upload_lattice(ctx, L):
    """Uploads the cells of lattice L; returns (cells, idx) with idx[v] the device node of L's v."""
    c = sh.lattice_cells(L)
    ctx.upload(c["sites"], c["held"], c["e"], c["g"], beams=c["beams"], spacing=c["spacing"])
    o = np.argmax(L.around[:, 1] != L.around[:, 0], axis=1)
    faces = np.concatenate([o[:, None], L.around[:, 1] >> L.level], axis=1).astype(np.int32)
    idx = ctx.lookup_faces(faces)
    assert (idx >= 0).all() and len(np.unique(idx)) == len(idx) == ctx.nodes
    return c, idx


def test_lattice_upload_solves_as_oracle(ctx):
    occ, theta = sh.generate(24, 24, 24, seed=8, sigma=0.5)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    L = orc.Lattice(1, sites.astype(np.int64) << 1, mats_from(e, g), held)
    c, idx = upload_lattice(ctx, L)
    gone = (L.joined_to_held() == 0).astype(np.uint8)
    j = ctx.joined_to_held()
    assert ((j[idx] == 0) == gone.astype(bool)).all()
    f = np.zeros((L.n_nodes, 6))
    f[:, 2] = -L.m * orc.GRAVITY
    x_or, _ = L.motions(f, gone=gone, tol=1e-13, max_iter=5000)
    fd = np.zeros((ctx.nodes, 6))
    fd[idx] = f
    for prec in (sh.FP64, sh.MIXED):
        x, st = ctx.solve(fd, precision=prec, tol=1e-10)
        assert st.status == 0 and rel(x[idx], x_or) < 1e-8


def _edited(n, seed):
    """A generated structure of strength 1 everywhere, made stronger by as_generated against
    itself, then with a block of cells cut out of its middle (an edit): the edited lattice made
    as strong as the generated one beam by beam."""
    occ, theta = sh.generate(n, n, n, seed=seed, sigma=0.5, pillar_spacing=6.0, pillar_min=2, pillar_max=4,
                             slab_min=5, slab_max=6)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    mats = mats_from(e, g, strength=(1.0, 1.0, 1.0))
    L = orc.Lattice(0, sites.astype(np.int64), mats, held)
    keep = ~((sites[:, 2] >= 2) & (sites[:, 2] <= 3) & (sites[:, 0] < n // 2))
    E = orc.Lattice(0, sites[keep].astype(np.int64), mats[keep], held[keep])
    E.as_generated(L, tol=1e-10)
    return E


def test_as_generated_used_and_settle(ctx):
    E = _edited(16, 5)
    c, idx = upload_lattice(ctx, E)
    assert len(np.unique(c["strength"][c["beams"] != 0])) > 3  # strengths differ beam by beam
    # bond used shares with the beams' own strengths
    rng = np.random.default_rng(5)
    x = rng.normal(size=(E.n_nodes, 6)) * 1e-2
    x[E.held.astype(bool)] = 0
    xd = np.zeros((ctx.nodes, 6))
    xd[idx] = x
    bob = c["beam_of_bond"]
    u_dev = ctx.bond_used(c["strength"], xd).reshape(-1, 4)[bob[0::4]].reshape(-1)
    u_or = E.used(E.moved(x))
    assert rel(u_dev, u_or) < 1e-12
    # settle: the same rounds, gone set and fallen share
    o = E.settle(tol=1e-10)
    f = np.zeros((ctx.nodes, 6))
    f[idx] = E.gravity()
    gone, beams, rounds, st = ctx.settle(c["strength"], f, precision=sh.MIXED, tol=1e-8, warm_start=True)
    assert st.status == 0
    gone = gone[idx]
    free = ~E.held.astype(bool)
    fallen = E.m[gone.astype(bool)].sum() / E.m[free].sum()
    print(f"as_generated + edit: {rounds} rounds (oracle {o['rounds']}), fallen {fallen:.4f}")
    assert rounds == o["rounds"] and o["rounds"] >= 2
    assert (gone == o["gone"]).all() and abs(fallen - o["fallen"]) <= 1e-12


def test_joined_edges_matches_oracle(ctx):
    """st_joined_edges (bodies::joined_to_held with the reference's own arguments: nodes, bonds,
    alive) against the oracle, bit for bit, with 60% of the beams dead (all four bonds; the odd
    dead bond alone rarely cuts anything off) and a few single bonds."""
    occ, theta = sh.generate(20, 20, 20, seed=12, sigma=0.5)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    L = orc.Lattice(0, sites.astype(np.int64), mats_from(e, g), held)
    rng = np.random.default_rng(12)
    alive = np.repeat(rng.random(L.n_bonds // 4) > 0.6, 4).astype(np.uint8)
    alive[rng.random(L.n_bonds) < 0.05] = 0
    want = L.joined_to_held(alive)
    got = ctx.joined_edges(L.held, L.ij, alive)
    assert 0 < want.sum() < L.n_nodes and (got == want).all()
    assert (ctx.joined_edges(L.held, L.ij) == L.joined_to_held()).all()
