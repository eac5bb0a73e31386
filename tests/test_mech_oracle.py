"""The oracle of mechanics::solve (oracle/mechanics_oracle.cpp) pinned without the reference:
the element's own test (mechanics/src/element.rs:164-187), the lift against explicit rigid
transfers, the inverses against numpy, the parts against a brute-force search, the
V-cycle's symmetry and positivity, and the solves against dense solves."""
import numpy as np
import pytest

import mech_harness as mh
import oracle as orc

H = 0.125


def skew_x(r):
    """X(r) w = w x r (mechanics/src/solve.rs:77-78)."""
    return np.array([[0.0, r[2], -r[1]], [-r[2], 0.0, r[0]], [r[1], -r[0], 0.0]])


def test_unit_cube_turned_bears_nothing_and_stretched_is_as_stiff_as_its_material():
    """element.rs:164-187 on the unit cube's stiffness: rigid motions make no forces; pulled
    along x by a thousandth, free across (Poisson's ratio 0.3), the far face's force over its
    area is the material's stiffness."""
    k = orc.mech_unit_k().transpose(0, 2, 1, 3).reshape(24, 24)
    home = mh.CORNERS.astype(float)
    for t in np.eye(3):
        assert np.abs(k @ np.tile(t, 8)).max() < 1e-14
    for w in np.eye(3):  # a small turn about the middle
        u = np.cross(w, home - 0.5)
        assert np.abs(k @ u.reshape(-1)).max() < 1e-14
    u = np.stack([home[:, 0] * 0.001, home[:, 1] * -0.3 * 0.001, home[:, 2] * -0.3 * 0.001], axis=1)
    f = (k @ u.reshape(-1)).reshape(8, 3)
    pull = f[mh.CORNERS[:, 0] == 1, 0].sum()
    assert abs(pull / 0.001 - 1.0) < 1e-6
    assert np.abs(k - k.T).max() < 1e-15 and np.linalg.eigvalsh(k)[6] > 0  # six rigid modes, no more


def test_system_builder_uses_the_oracles_unit_cube():
    """The benchmark's system builder (parallax-world_amd/mech_systems.py) has its own unit
    cube stiffness: it is the oracle's."""
    assert np.abs(mh.unit_k() - orc.mech_unit_k()).max() < 1e-15


def test_lift_is_the_rigid_transfer():
    rng = np.random.default_rng(0)
    for _ in range(5):
        m = rng.normal(size=(6, 6))
        r, s = rng.normal(size=3), rng.normal(size=3)
        t = lambda v: np.block([[np.eye(3), skew_x(v)], [np.zeros((3, 3)), np.eye(3)]])
        want = t(r).T @ m @ t(s)
        assert np.abs(orc.mech_lift(m, r, s) - want).max() < 1e-12 * np.abs(want).max()


def test_inverse_with_dropped_pivots():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(6, 6))
    m = a @ a.T + 6 * np.eye(6)
    assert np.abs(orc.mech_inverse6(m) @ m - np.eye(6)).max() < 1e-12
    # a direction the matrix does not reach stays still: its row and column of the inverse are 0
    m[5, :] = m[:, 5] = 0.0
    inv = orc.mech_inverse6(m)
    assert (inv[5] == 0).all() and (inv[:, 5] == 0).all()
    assert np.abs(inv[:5, :5] @ m[:5, :5] - np.eye(5)).max() < 1e-12


SYSTEMS = {
    "post": lambda: mh.system(*mh.post(8), h=H),
    "turned blob": lambda: mh.system(*mh.blob(8, 0.6, 1), h=H, turn=0.4, seed=2),
    "two posts apart": lambda: mh.system(np.concatenate([mh.post(5)[0], mh.post(5)[0] + [4, 0, 0]]),
                                         lambda k: k[:, 1] == 0, h=H),
}


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_solve_is_the_dense_solve(name):
    s, b, _, _ = SYSTEMS[name]()
    a = s.dense()
    assert np.abs(a - a.T).max() <= 1e-12 * np.abs(a).max() and np.linalg.eigvalsh(a)[0] > 0
    x_dense = np.linalg.solve(a, b.reshape(-1)).reshape(-1, 3)
    assert np.abs(s.times(x_dense) - b).max() < 1e-9 * np.abs(b).max()  # System::times is the matrix
    o = orc.MechSolver(s, H)
    x, rounds = o.solve(s, b, tol=1e-12)
    assert rounds is not None and rounds < 300
    assert np.abs(x - x_dense).max() < 1e-9 * np.abs(x_dense).max()


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_vcycle_is_symmetric_and_positive(name):
    """The preconditioner is a linear map P, symmetric and positive (what makes CG work,
    solve.rs:303-308): y^T P z = z^T P y and y^T P y > 0."""
    s, b, _, _ = SYSTEMS[name]()
    o = orc.MechSolver(s, H)
    rng = np.random.default_rng(3)
    y, z = rng.normal(size=b.shape), rng.normal(size=b.shape)
    py, pz = o.precondition(s, y), o.precondition(s, z)
    assert abs((y * pz).sum() - (z * py).sum()) <= 1e-10 * abs((y * py).sum())
    assert (y * py).sum() > 0 and (z * pz).sum() > 0
    assert np.abs(o.precondition(s, 2 * y + z) - (2 * py + pz)).max() <= 1e-10 * np.abs(py).max()


def test_parts_are_the_joined_pieces_of_each_octree_cell():
    """Level 1's parts against a breadth-first search: two unknowns share a part when a chain
    of pairs joins them without leaving their octree cell a level up; parts are numbered in
    the order their first unknown comes (solve.rs:210-250)."""
    s, b, _, _ = SYSTEMS["turned blob"]()
    o = orc.MechSolver(s, H)
    of, r = o.level_map(0, s.n)
    parent = np.floor_divide(s.sites, 2)
    adj = [[] for _ in range(s.n)]
    for i, j in s.ij:
        if (parent[i] == parent[j]).all():
            adj[i].append(j)
            adj[j].append(i)
    want = np.full(s.n, -1)
    count = 0
    for i in range(s.n):
        if want[i] >= 0:
            continue
        want[i] = count
        stack = [i]
        while stack:
            v = stack.pop()
            for w in adj[v]:
                if want[w] < 0:
                    want[w] = count
                    stack.append(w)
        count += 1
    assert (of == want).all()
    # each part's middle is its members' mean: the offsets sum to nothing
    sums = np.zeros((count, 3))
    np.add.at(sums, of, r)
    assert np.abs(sums).max() < 1e-12
    levels, _ = o.levels()
    assert levels[0][0] == count
