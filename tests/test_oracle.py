"""The oracle pinned: every check here is independent of the restatement's own formulas.

- reference-held pins: the cells_nest test (statics/src/lattice.rs:169-181), the node-mass
  conservation of gather.rs:400-412 applied to build, and the level-0 "used" column of the
  edits table in docs/statics.md:236-245 (the only published numbers on this path);
- closed forms: the SECTION identity (statics/src/lattice.rs:21-25), rigid motions, the
  principal stresses of Bond::used;
- brute force: dense K assembled here from the bond list, solved by numpy;
- independent restatements: BFS for joined_to_held, a triple loop for hold_bulk.
"""
import itertools

import numpy as np
import pytest

import oracle as orc
import scenes

S0 = orc.UNIT


def lattice_of(sites, mats, held, level=0):
    return orc.Lattice(level, sites, mats, held)


# ---------------------------------------------------------------- reference pins

def test_cells_nest():
    """statics/src/lattice.rs:169-181: coarse_site nests, negative coordinates included."""
    for p in ([0, 0, 0], [1, -1, 5], [-3, 7, -8]):
        for l in range(4):
            up = orc.coarse_site(p, l + 1)
            assert (orc.coarse_site(orc.coarse_site(p, l), l + 1) == up).all()
            for k in range(3):
                assert up[k] <= p[k] < up[k] + (1 << (l + 1))


def test_mass_is_kept():
    """gather.rs:400-412's pin, on build: the nodes' mass is the free cells' mass."""
    sites, mats, held = scenes.block(3, 4, 2)
    L = lattice_of(sites, mats, held)
    free = held == 0
    assert abs(L.m.sum() - (mats[free, 0] * S0 ** 3).sum()) <= 1e-12 * L.m.sum()


@pytest.mark.parametrize("name,scene,want,tol", [
    ("a stone beam made 4 m", scenes.cantilever(orc.STONE, 0.5, 4.0), 0.43, 0.006),
    ("made 6 m", scenes.cantilever(orc.STONE, 0.5, 6.0), 0.97, 0.006),
    ("a slab, one column cut", scenes.portico(1, 0.5), 1.03, 0.006),
    ("two columns cut", scenes.portico(2, 0.5), 7.3, 0.06),
    ("a tree cut 30 %", scenes.tree(8.0, -1.0, 0.3), 0.37, 0.006),
    ("cut 70 %", scenes.tree(8.0, -1.0, 0.7), 3.3, 0.06),
])
def test_documented_used(name, scene, want, tol):
    """docs/statics.md:236-245, column "Used at 0": the most any bond uses, level 0, as built."""
    sites, mats, held = scenes.solids(0, *scene)
    L = lattice_of(sites, mats, held)
    u = L.used_standing(tol=1e-6)
    assert abs(u.max() - want) <= tol, (name, u.max())


# ---------------------------------------------------------------- closed forms

def test_section_identity():
    """A block strained along x with its sides held carries s^2 = A180 + sqrt(2)(1 + G/E) A90 of E
    (statics/src/lattice.rs:21-25): the virial stress of the middle cell's bonds."""
    sites, mats, held = scenes.block(3, 3, 3, held_layer=False)
    L = lattice_of(sites, mats, held)
    eps = 1e-4
    x = np.zeros((L.n_nodes, 6))
    x[:, 0] = eps * L.x[:, 0]
    f = L.moved(x)
    mid = int(np.flatnonzero((sites == [1, 1, 1]).all(axis=1))[0])
    bonds = np.flatnonzero(np.repeat(L.beam_cell, 4) == mid)
    lever = L.x[L.ij[bonds, 1], 0] - L.x[L.ij[bonds, 0], 0]
    sigma = (f[bonds, 0] * lever).sum() / S0 ** 3
    e, g = orc.STONE[1], orc.STONE[2]
    want = e * eps * (0.336 + np.sqrt(2.0) * (1.0 + g / e) * 0.336)
    assert abs(sigma / want - 1.0) < 1e-12
    assert abs(want / (e * eps) - 1.0012) < 1e-4  # "as stiff as its material", to about 0.12%


def test_rigid_motion_moves_no_bond():
    sites, mats, held = scenes.block(2, 2, 2, held_layer=False)
    L = lattice_of(sites, mats, held)
    rng = np.random.default_rng(0)
    a, w = rng.normal(size=3), rng.normal(size=3)
    x = np.zeros((L.n_nodes, 6))
    x[:, :3] = a + np.cross(w, L.x)
    x[:, 3:] = w
    f = L.moved(x)
    scale = orc.STONE[1] * L.a.max() / L.d.min() * (abs(a).max() + abs(w).max())
    assert np.abs(f).max() <= 1e-12 * scale


def test_bond_used_is_principal_stress():
    """model.rs:123-129: s1, s3 are the principal stresses of a pull sigma and a shear tau."""
    sites, mats, held = scenes.block(1, 1, 1, held_layer=False)
    L = lattice_of(sites, mats, held)
    rng = np.random.default_rng(1)
    f = rng.normal(size=(L.n_bonds, 3)) * 1e5
    strength = np.tile([5e6, 100e6, 10e6], (L.n_bonds, 1))
    u = L.used(f, strength)
    for b in range(0, L.n_bonds, 7):
        t = f[b] / L.a[b]
        sig = t @ L.n[b]
        tau = np.sqrt(max(t @ t - sig * sig, 0.0))
        s3, s1 = np.linalg.eigvalsh([[sig, tau], [tau, 0.0]])
        want = max(s1 / 5e6, -s3 / 100e6, (s1 - s3) / (2 * 10e6))
        assert abs(u[b] - want) <= 1e-12 * max(1.0, abs(want))


def test_invert6():
    rng = np.random.default_rng(2)
    for _ in range(20):
        a = rng.normal(size=(6, 6))
        a = a @ a.T + 6 * np.eye(6)
        assert np.abs(orc.invert6(a) @ a - np.eye(6)).max() < 1e-12
    # ties in the pivot column (the reference keeps the last maximum)
    a = np.eye(6)[::-1] * 2.0 + np.triu(np.ones((6, 6)), 1) * 0.0
    a[0, 0] = a[5, 0] = 2.0
    a[5, 5] = 1.0
    assert np.abs(orc.invert6(a) @ a - np.eye(6)).max() < 1e-12


# ---------------------------------------------------------------- brute force

def dense_k(L, alive=None, gone=None):
    """K over the free, not gone nodes, assembled here bond by bond from the bond list."""
    n = L.n_nodes
    alive = np.ones(L.n_bonds, bool) if alive is None else alive.astype(bool)
    gone = np.zeros(n, bool) if gone is None else gone.astype(bool)
    part = (~L.held.astype(bool)) & ~gone
    idx = -np.ones(n, int)
    idx[part] = np.arange(part.sum())
    K = np.zeros((6 * part.sum(), 6 * part.sum()))
    for b in range(L.n_bonds):
        i, j = L.ij[b]
        if not alive[b] or gone[i] or gone[j]:
            continue
        nn = np.outer(L.n[b], L.n[b])
        k = L.a[b] / L.d[b]
        S = L.e[b] * k * nn + L.g[b] * k * (np.eye(3) - nn)

        def T(v):
            r = L.p[b] - L.x[v]
            t = np.zeros((3, 6))
            t[:, :3] = np.eye(3)
            t[:, 3:] = -np.array([[0, -r[2], r[1]], [r[2], 0, -r[0]], [-r[1], r[0], 0]])
            return t  # T x = u + w x r

        D = np.zeros((3, 6 * part.sum()))
        if idx[j] >= 0:
            D[:, 6 * idx[j]:6 * idx[j] + 6] += T(j)
        if idx[i] >= 0:
            D[:, 6 * idx[i]:6 * idx[i] + 6] -= T(i)
        K += D.T @ S @ D
    return K, part


@pytest.mark.parametrize("cells", [
    [[0, 0, 0]],
    [[0, 0, 0], [1, 0, 0]],
    [[0, 0, 0], [0, 1, 0], [0, 1, 1], [1, 1, 1]],
    [[x, y, z] for x in range(2) for y in range(2) for z in range(2)],
])
def test_motions_brute_force(cells):
    sites = np.array([[c[0], c[1] + 1, c[2]] for c in cells] + [[x, 0, z] for x in range(-1, 3) for z in range(-1, 3)])
    held = np.array([0] * len(cells) + [1] * 16, np.uint8)
    mats = np.tile(np.array(orc.STONE), (len(sites), 1))
    L = lattice_of(sites, mats, held)
    K, part = dense_k(L)
    assert np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
    assert np.linalg.eigvalsh(K).min() > 0
    f = L.gravity()
    x_brute = np.linalg.solve(K, f[part].ravel()).reshape(-1, 6)
    x, st = L.motions(f, tol=1e-14, max_iter=2000)
    assert np.abs(x[part] - x_brute).max() <= 1e-9 * np.abs(x_brute).max()
    assert np.abs(x[~part]).max() == 0.0
    # the restatement's Level::apply is the same K
    y = L.apply(x)
    assert np.abs(y[part].ravel() - K @ x[part].ravel()).max() <= 1e-9 * np.abs(K @ x[part].ravel()).max()


def test_equilibrium_and_compliance():
    """Sum of bond forces at the held nodes balances the loads (force and moment); f.x = x.K.x."""
    sites, mats, held = scenes.block(3, 3, 2)
    L = lattice_of(sites, mats, held)
    f = L.gravity()
    x, _ = L.motions(f, tol=1e-13, max_iter=5000)
    fb = L.moved(x)
    heldn = L.held.astype(bool)
    # K x = f with K x at i = -sum T_i^T f_b: a bond pulls its i node by +f_b, its j node by -f_b
    R = np.zeros((L.n_nodes, 3))
    M = np.zeros((L.n_nodes, 3))
    np.add.at(R, L.ij[:, 0], fb)
    np.add.at(R, L.ij[:, 1], -fb)
    np.add.at(M, L.ij[:, 0], np.cross(L.p, fb))
    np.add.at(M, L.ij[:, 1], -np.cross(L.p, fb))
    # motions reads loads on free nodes only; the bonds' pull on the held nodes carries them all
    total_load = f[~heldn, :3].sum(axis=0)
    load_moment = np.cross(L.x[~heldn], f[~heldn, :3]).sum(axis=0)
    assert np.abs(R[heldn].sum(axis=0) - total_load).max() <= 1e-8 * np.abs(total_load).max()
    assert np.abs(M[heldn].sum(axis=0) - load_moment).max() <= 1e-8 * np.abs(load_moment).max()
    y = L.apply(x)
    assert abs((f * x).sum() - (x * y).sum()) <= 1e-10 * abs((f * x).sum())


# ---------------------------------------------------------------- independent restatements

def bfs_joined(L, alive):
    adj = [[] for _ in range(L.n_nodes)]
    for b in np.flatnonzero(alive):
        i, j = L.ij[b]
        adj[i].append(j)
        adj[j].append(i)
    seen = L.held.astype(bool).copy()
    queue = list(np.flatnonzero(seen))
    while queue:
        v = queue.pop()
        for w in adj[v]:
            if not seen[w]:
                seen[w] = True
                queue.append(w)
    return seen


def test_joined_to_held_bfs():
    import statics_hip as sh
    occ, theta = sh.generate(16, 16, 16, seed=3)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    mats = np.tile(np.array(orc.STONE), (len(sites), 1))
    L = lattice_of(sites, mats, held)
    rng = np.random.default_rng(3)
    for p_dead in (0.0, 0.3):
        beam_alive = rng.random(L.n_bonds // 4) >= p_dead
        alive = np.repeat(beam_alive, 4).astype(np.uint8)
        assert (L.joined_to_held(alive).astype(bool) == bfs_joined(L, alive)).all()


def test_hold_bulk_triple_loop():
    rng = np.random.default_rng(4)
    cells = {(x, y, z) for x in range(8) for y in range(8) for z in range(8) if rng.random() < 0.8}
    sites = np.array(sorted(cells), np.int64)
    for level, radius in ((0, 0.2), (0, 0.3), (1, 0.5)):
        s = S0 * (1 << level)
        sc = sites << level
        have = {tuple(c) for c in sc}
        r = int(np.ceil(radius / s))
        near = [(a << level, b << level, c << level) for a, b, c in itertools.product(range(-r, r + 1), repeat=3)
                if np.sqrt(float(a * a + b * b + c * c)) * s <= radius]
        want = [all((c[0] + o[0], c[1] + o[1], c[2] + o[2]) in have for o in near) for c in sc]
        got = orc.hold_bulk(level, sc, radius).astype(bool)
        assert (got == np.array(want)).all()
        assert got.any() or radius > 0.25


def test_settle_cantilever_against_beam_theory():
    """docs/statics.md:44-46: a beam out of a wall holds to 0.9-1.1 of beam theory's length,
    sqrt(f h / 3 rho g), f the least of ft, fc, 2 fs (statics/examples/statics.rs:146-163)."""
    mat, h = orc.STONE, 0.25
    f = min(mat[3], mat[4], 2 * mat[5])
    theory = np.sqrt(f * h / (3 * mat[0] * orc.GRAVITY))
    lo, hi = 1, int(2 * theory / S0)
    while hi - lo > 1:
        mid = (lo + hi) // 2
        sites, mats, held = scenes.solids(0, *scenes.cantilever(mat, h, mid * S0))
        L = lattice_of(sites, mats, held)
        if L.settle(tol=1e-6)["rounds"] > 1:
            hi = mid
        else:
            lo = mid
    assert 0.9 <= lo * S0 / theory <= 1.1


# ---------------------------------------------------------------- gather.rs


def test_gather_bar_weighs_as_much():
    """statics/src/gather.rs:400-412: a steel bar keeps its weight through three gatherings."""
    fill = lambda x: (orc.STEEL, False) if (abs(x[0]) < 4.0 and abs(x[1]) < 0.2 and abs(x[2]) < 0.2) else None
    lat = orc.Lattice(0, *scenes.solids(0, [-4.5, -1.0, -1.0], [4.5, 1.0, 1.0], fill))
    whole = lat.m.sum()
    for level in (1, 2, 3):
        lat = lat.gathered()
        assert lat.level == level and lat.n_bonds > 0
        assert abs(lat.m.sum() / whole - 1.0) < 1e-9


def _centre_beams(lat):
    """The beams of the coarse cell at the origin: (class, area, ft) of each."""
    s = orc.spacing(lat.level)
    c = (np.zeros(3) + 0.5 * s)
    inc = lambda v: (lat.around[v][0] == 0).all() or (lat.around[v][1] == 0).all()
    out = []
    for q in range(lat.n_bonds // 4):
        i, j = lat.ij[4 * q]
        if inc(i) and inc(j):
            u, v = lat.x[i] - c, lat.x[j] - c
            k = int(round(1.0 + u @ v / (np.linalg.norm(u) * np.linalg.norm(v))))
            out.append((k, 4.0 * lat.a[4 * q], lat.bond_mat[4 * q, 3]))
    return out


@pytest.mark.parametrize("mat", [orc.STONE, orc.WOOD])
def test_gather_full_cell_is_section(mat):
    """A full cell gathered against a full cell's measure is SECTION of its material: all 15
    beams (3 between opposite faces, 12 between neighbouring ones) of section SECTION s^2
    (statics/src/lattice.rs:20-25) and the material's own strength (gather.rs:187-192)."""
    s = orc.spacing(1)
    lat = orc.Lattice(0, *scenes.solids(0, [-1.5 * s] * 3, [1.5 * s] * 3, lambda x: (mat, False)))
    beams = _centre_beams(lat.gathered())
    assert sorted(k for k, _, _ in beams) == [0] * 3 + [1] * 12
    for _, area, ft in beams:
        assert area == pytest.approx(0.336 * s * s, rel=1e-12)
        assert ft == pytest.approx(mat[3], rel=1e-12)
    # measured as it is, the same cell's beams differ by class but not within one
    raw = _centre_beams(lat.gathered(against_full=False))
    for k in (0, 1):
        a = [area for kk, area, _ in raw if kk == k]
        assert max(a) == pytest.approx(min(a), rel=1e-9)


def test_gather_truss_stays_truss():
    """Matter that only lies near other matter in a cell does not join it (gather.rs:14-15):
    two level-0 cells touching at a corner inside one coarse cell give two separate sets of
    beams, three among each one's own three handles and none across."""
    lat = orc.Lattice(0, np.array([[0, 0, 0], [1, 1, 1]]), [orc.STONE] * 2, np.zeros(2, np.uint8))
    g = lat.gathered()
    assert g.n_nodes == 6 and g.n_bonds == 6 * 4
    low = lambda v: int(g.around[v].min() < 0)  # a face on the coarse cell's lower side
    for q in range(6):
        i, j = g.ij[4 * q]
        assert low(i) == low(j)
    assert g.m.sum() == pytest.approx(lat.m.sum(), rel=1e-12)


def test_gather_solve():
    """gather.rs:165-186: numpy's solution for a regular system; a direction it cannot tell
    (a zero row and column, last: the pivot search keeps the last of equal maxima, so a zero
    column earlier would swap a later row in) solved as zero, the rest as without it."""
    rng = np.random.default_rng(5)
    n, rhs = 12, 3
    a = rng.standard_normal((n, n)) + n * np.eye(n)
    b = rng.standard_normal((n, rhs))
    x = orc.gather_solve(np.hstack([a, b]), rhs)
    assert np.allclose(x, np.linalg.solve(a, b), rtol=1e-12, atol=1e-12)
    a[n - 1, :] = 0.0
    a[:, n - 1] = 0.0
    b[n - 1] = 0.0
    x = orc.gather_solve(np.hstack([a, b]), rhs)
    keep = list(range(n - 1))
    assert (x[n - 1] == 0.0).all()
    assert np.allclose(x[keep], np.linalg.solve(a[np.ix_(keep, keep)], b[keep]), rtol=1e-12, atol=1e-12)


def _units(at, pairs):
    n = 3 * len(at)
    out = []
    for a, b in pairs:
        e = (at[b] - at[a]) / np.linalg.norm(at[b] - at[a])
        u = np.zeros((n, n))
        ee = np.outer(e, e)
        u[3 * a:3 * a + 3, 3 * a:3 * a + 3] += ee
        u[3 * b:3 * b + 3, 3 * b:3 * b + 3] += ee
        u[3 * a:3 * a + 3, 3 * b:3 * b + 3] -= ee
        u[3 * b:3 * b + 3, 3 * a:3 * a + 3] -= ee
        out.append(u)
    return out


def test_gather_beams_fit():
    """gather.rs:116-161: a stiffness made of beams is fitted by those beams; a beam asked to
    be negative is dropped and the rest refitted (numpy's least squares on what is kept);
    pairs not joined get none."""
    at = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]], float) * 0.25
    pairs = [(a, b) for a in range(6) for b in range(a + 1, 6)]
    u = _units(at, pairs)
    want = {(0, 1): 3.0, (0, 2): 1.5, (2, 5): 0.7, (3, 4): 2.2}
    k = sum(v * u[pairs.index(p)] for p, v in want.items())
    joined = np.ones((6, 6), np.uint8)
    got = orc.beams(at, k, joined)
    big = {(a, b): v for a, b, v in got if v > 1e-9}
    assert set(big) == set(want)
    for p, v in want.items():
        assert big[p] == pytest.approx(v, rel=1e-9)
    # one beam negative: dropped, the rest is numpy's least squares over the kept pairs
    k2 = k - 2.0 * u[pairs.index((1, 3))]
    got = orc.beams(at, k2, joined)
    assert all(v > 0 for _, _, v in got) and (1, 3) not in [(a, b) for a, b, _ in got]
    kept = [p for p in pairs if p in [(a, b) for a, b, _ in got]]
    A = np.stack([u[pairs.index(p)].ravel() for p in kept], 1)
    lsq = np.linalg.lstsq(A, k2.ravel(), rcond=None)[0]
    assert np.allclose([v for _, _, v in got], lsq, rtol=1e-9, atol=1e-12)
    # not joined: no beam between 0 and 1
    joined[0, 1] = joined[1, 0] = 0
    assert (0, 1) not in [(a, b) for a, b, _ in orc.beams(at, k, joined)]


def _blob(n, seed):
    rng = np.random.default_rng(seed)
    sites, mats, held = [], [], []
    for x in range(n):
        for y in range(n):
            for z in range(n):
                if rng.random() < 0.7 or y == 0:
                    sites.append((x - n // 2, y, z - n // 2))
                    mats.append(orc.STONE if rng.random() < 0.6 else orc.WOOD)
                    held.append(1 if y == 0 else 0)
    return orc.Lattice(0, np.array(sites), np.array(mats), np.array(held, np.uint8))


@pytest.mark.parametrize("levels", [0, 1])
def test_gather_cell_stiffness_invariants(levels):
    """gather.rs:59-114, 291-303: a cell's stiffness between its handles is its matter's
    stiffness condensed onto them, so it is symmetric, positive semi-definite, and moving every
    handle by one translation (the free parts following) asks no force. A wrong sign, a dropped
    term or a transposed product in the elimination, the handle columns or the reactions breaks
    one of these; a full cell's normalisation cannot hide it."""
    lat = _blob(8, 11)
    for _ in range(levels):
        lat = lat.gathered()
    ks = lat.cell_stiffness()
    assert len(ks) >= 8
    for k in ks:
        m = len(k)
        scale = np.abs(k).max()
        assert scale > 0
        assert np.abs(k - k.T).max() <= 1e-9 * scale
        for t in np.eye(3):
            assert np.abs(k @ np.tile(t, m // 3)).max() <= 1e-9 * scale * m
        ev = np.linalg.eigvalsh(0.5 * (k + k.T))
        assert ev.min() >= -1e-9 * ev.max()
        assert (ev > 1e-6 * ev.max()).sum() >= 1


def test_gather_energy_is_compliance():
    """Bond::energy (model.rs:132-136), which weighs the finer bonds' shares of a pulled beam,
    is half the force through the inverse of the bond's own spring matrix."""
    lat = _blob(4, 2)
    rng = np.random.default_rng(3)
    f = rng.standard_normal((lat.n_bonds, 3)) * 1e3
    w = lat.energy(f)
    for b in rng.choice(lat.n_bonds, 50, replace=False):
        k = lat.a[b] / lat.d[b]
        nn = np.outer(lat.n[b], lat.n[b])
        s = lat.e[b] * k * nn + lat.g[b] * k * (np.eye(3) - nn)
        assert w[b] == pytest.approx(0.5 * f[b] @ np.linalg.solve(s, f[b]), rel=1e-12)
