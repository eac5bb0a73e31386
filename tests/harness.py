"""Lining the device's node and beam order up with the oracle's (the reference's)."""
import numpy as np

import oracle as orc
import statics_hip as sh

FACES = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]])
PAIRS = [(a, b) for a in range(6) for b in range(a + 1, 6)]
PAIR_OF = {p: q for q, p in enumerate(PAIRS)}


def mats_from(e, g, rho=1.0, strength=(1.0, 1.0, 1.0)):
    m = np.zeros((len(e), 6))
    m[:, 0] = rho
    m[:, 1] = e
    m[:, 2] = g
    m[:, 3:] = strength
    return m


class Pair:
    """The same cells as an oracle Lattice and as an uploaded device context."""

    def __init__(self, ctx, sites, mats, held, level=0):
        self.sites = np.ascontiguousarray(sites, np.int64)
        self.mats = np.ascontiguousarray(mats, np.float64)
        self.held_cells = np.ascontiguousarray(held, np.uint8)
        self.level = level
        self.L = orc.Lattice(level, self.sites, self.mats, self.held_cells)
        self.ctx = ctx
        ctx.upload((self.sites >> level).astype(np.int32), self.held_cells, self.mats[:, 1], self.mats[:, 2],
                   spacing=orc.UNIT * (1 << level))
        L = self.L
        # node v of the oracle is the face (o, greater cell) on the device
        o = np.argmax(L.around[:, 1] != L.around[:, 0], axis=1)
        faces = np.concatenate([o[:, None], L.around[:, 1] >> level], axis=1).astype(np.int32)
        self.idx = ctx.lookup_faces(faces)
        assert (self.idx >= 0).all() and len(np.unique(self.idx)) == len(self.idx) == ctx.nodes
        # each oracle beam: its cell (index into sites) and pair q
        nb = L.n_bonds // 4
        self.beam_cell = L.beam_cell
        i, j = L.ij[0::4, 0], L.ij[0::4, 1]
        site = self.sites[self.beam_cell]

        def face_of(v):
            a = L.around[v]
            other = np.where((a[:, 0] == site).all(axis=1)[:, None], a[:, 1], a[:, 0])
            d = (other - site) >> level
            return np.array([int(np.flatnonzero((FACES == dd).all(axis=1))[0]) for dd in d])

        fa, fb = face_of(i), face_of(j)
        self.beam_q = np.array([PAIR_OF[(a, b)] for a, b in zip(fa, fb)])
        assert len(self.beam_q) == nb

    def to_dev(self, a):
        out = np.zeros((self.ctx.nodes,) + a.shape[1:])
        out[self.idx] = a
        return out

    def from_dev(self, a):
        return a[self.idx]

    def beam_masks(self, alive_bonds):
        """Per cell 15-bit alive masks from the oracle's per-bond alive array (beams break whole)."""
        m = np.full(len(self.sites), 0x7FFF, np.uint16)
        dead = ~alive_bonds.reshape(-1, 4).all(axis=1)
        for b in np.flatnonzero(dead):
            m[self.beam_cell[b]] &= ~np.uint16(1 << self.beam_q[b])
        return m

    def bond_forces_dev(self, x_oracle_order):
        """Device bond forces laid out as the oracle's bonds."""
        f = self.ctx.bond_forces(self.to_dev(x_oracle_order))  # [cell][q][k][3]
        return f[self.beam_cell, self.beam_q].reshape(-1, 3)


def generated(ctx, n, seed, sigma=0.5, **kw):
    occ, theta = sh.generate(n, n, n, seed=seed, sigma=sigma, **kw)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    return Pair(ctx, sites, mats_from(e, g), held)


def point_loads(P, gone, rng, n_points, gravity=True):
    """The configurations' loads in the oracle's node order: gravity of weight 1 per free cell,
    a sixth on each of its faces, along -z, plus n_points point loads of magnitude U(1, 10) in
    uniformly random directions on the translations of distinct free nodes (not held, not
    gone). The reference takes loads through motions' closure, only for free nodes
    (statics/src/bodies.rs:429, 462)."""
    L = P.L
    f = np.zeros((L.n_nodes, 6))
    if gravity:
        f[:, 2] = -L.m / orc.UNIT ** 3
    free = np.flatnonzero((L.held == 0) & (gone == 0))
    pick = rng.choice(free, size=min(n_points, len(free)), replace=False)
    d = rng.normal(size=(len(pick), 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    f[pick, :3] += d * rng.uniform(1.0, 10.0, size=(len(pick), 1))
    return f


def device_loads(ctx, occ, joined, rng, n_points):
    """The same loads for a structure uploaded from occ (no oracle lattice), in the device's node
    order: each face's weight is a sixth of each free cell beside it."""
    faces = ctx.node_faces()
    nheld = ctx.node_held().astype(bool)
    o = faces[:, 0]
    plus = faces[:, 1:4]
    minus = plus - np.eye(3, dtype=np.int32)[o]
    n = np.array(occ.shape)

    def free_cell(c):
        ok = np.all((c >= 0) & (c < n), axis=1)
        cc = np.clip(c, 0, n - 1)
        return ok & (occ[cc[:, 0], cc[:, 1], cc[:, 2]] == 1)

    f = np.zeros((len(faces), 6))
    f[:, 2] = -(free_cell(plus).astype(np.float64) + free_cell(minus)) / 6.0
    free = np.flatnonzero(joined.astype(bool) & ~nheld)
    pick = rng.choice(free, size=min(n_points, len(free)), replace=False)
    d = rng.normal(size=(len(pick), 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    f[pick, :3] += d * rng.uniform(1.0, 10.0, size=(len(pick), 1))
    f[nheld | ~joined.astype(bool)] = 0.0
    return f
