"""Systems for mechanics::solve as the live engine makes them (mechanics/src/lib.rs:638-676):
cells of an octree level as co-rotated eight-node cubes (mechanics/src/element.rs), their
corners welded into nodes, the nodes a held site holds left out, and

    (M + (dt^2 + DAMPING dt) K) v' = M v + dt F

over the rest: each node's mass (an eighth of each of its cells', lib.rs:391) on its
diagonal as block([1, 0, 0], m, m) = m I (lib.rs:202-204, 643), each cell's turned blocks
R K_ac R^T times E h (lib.rs:541-543, 650-672), one pair per two nodes of a cell (the
`pattern`, lib.rs:401-422: (p.min(q), p.max(q)) as first met, the block transposed where
p > q). Contacts are left out. The right-hand side is M v + dt M g with v small and random
(the cells at rest: their forces are nothing)."""
import numpy as np

import oracle as orc

CORNERS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]])
DAMPING = 0.3  # mechanics/src/lib.rs:63
GRAVITY = 9.81
WOOD = (600.0, 10e9)  # rho, e (mechanics/src/lib.rs:35)


def quat_matrix(q):
    """element.rs:89-95"""
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def system(sites, held_node, h=0.125, mat=WOOD, fill=None, turn=0.0, seed=0, dt=1.0 / 60.0):
    """The system of cells at `sites` (n, 3) int: held_node(corner coordinates (m, 3)) -> bool
    mask of the nodes a held site holds; `turn`: the cells' turns, radians at most (random
    axes); fill: each cell's share of its cube that bears (default 1). Returns (MechSystem,
    b, node count, the unknowns' corner coordinates)."""
    rng = np.random.default_rng(seed)
    sites = np.asarray(sites, np.int64)
    nc = len(sites)
    corners = (sites[:, None, :] + CORNERS[None]).reshape(-1, 3)
    keys, node_of = np.unique(corners, axis=0, return_inverse=True)
    node_of = node_of.reshape(nc, 8)
    nn = len(keys)
    mass = np.zeros(nn)
    rho, e = mat
    np.add.at(mass, node_of.reshape(-1), rho * h ** 3 / 8.0)
    held = np.asarray(held_node(keys), bool)
    free = np.full(nn, -1, np.int64)
    free[~held] = np.arange(int((~held).sum()))
    n = int((~held).sum())
    unit = orc.mech_unit_k()
    stiff = dt * dt + DAMPING * dt
    fill = np.ones(nc) if fill is None else np.asarray(fill, np.float64)
    # each cell's turn
    if turn > 0:
        ax = rng.normal(size=(nc, 3))
        ax /= np.linalg.norm(ax, axis=1, keepdims=True)
        ang = rng.uniform(0, turn, nc)
        q = np.concatenate([np.cos(ang / 2)[:, None], ax * np.sin(ang / 2)[:, None]], axis=1)
        R = np.array([quat_matrix(qq) for qq in q])
    else:
        R = np.broadcast_to(np.eye(3), (nc, 3, 3))
    # element stiffness R K_ac R^T for all cells: (nc, 8, 8, 3, 3)
    K = np.einsum("nip,acpq,njq->nacij", R, unit, R)
    kcell = e * fill * h * stiff
    diag = mass[~held][:, None, None] * np.eye(3)[None]
    fa = free[node_of]  # (nc, 8)
    for a in range(8):
        ok = fa[:, a] >= 0
        np.add.at(diag, fa[ok, a], K[ok, a, a] * kcell[ok, None, None])
    # pattern: pairs as first met, cell by cell, a < c
    index, ij, blocks = {}, [], []
    for i in range(nc):
        for a in range(8):
            p = fa[i, a]
            if p < 0:
                continue
            for c in range(a + 1, 8):
                q_ = fa[i, c]
                if q_ < 0:
                    continue
                key = (min(p, q_), max(p, q_))
                k = index.get(key)
                if k is None:
                    k = index[key] = len(ij)
                    ij.append(key)
                    blocks.append(np.zeros((3, 3)))
                m = K[i, a, c] * kcell[i]
                blocks[k] += m.T if p > q_ else m
    s = orc.MechSystem(keys[~held], keys[~held] * h, diag, np.array(ij, np.uint32).reshape(-1, 2),
                       np.array(blocks).reshape(-1, 3, 3))
    v = rng.normal(size=(n, 3)) * 1e-2
    b = (v + np.array([0.0, -GRAVITY * dt, 0.0])) * mass[~held][:, None]
    return s, b, nn, keys[~held]


def post(height, width=2, foot_held=True):
    """A post of width^2 x height cells standing on a held ground layer (the nodes at y = 0 held)."""
    sites = np.array([(x, y, z) for x in range(width) for y in range(height) for z in range(width)])
    return sites, (lambda k: k[:, 1] == 0) if foot_held else (lambda k: np.zeros(len(k), bool))


def blob(n, fill=0.5, seed=0):
    """Cells of an n^3 box kept with probability `fill`, plus a full bottom layer; nodes at y = 0 held."""
    rng = np.random.default_rng(seed)
    g = np.argwhere(np.ones((n, n, n), bool))
    keep = (rng.random(len(g)) < fill) | (g[:, 1] == 0)
    return g[keep], (lambda k: k[:, 1] == 0)
