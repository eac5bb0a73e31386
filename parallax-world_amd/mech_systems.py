"""Systems for mechanics::solve as the live engine makes them (mechanics/src/lib.rs:638-676),
for the benchmark and the tests: cells of an octree level as co-rotated eight-node cubes
(mechanics/src/element.rs), their corners welded into nodes, the nodes a held site holds left
out, and

    (M + (dt^2 + DAMPING dt) K) v' = M v + dt F

over the rest: each node's mass (an eighth of each of its cells', lib.rs:391) on its
diagonal as block([1, 0, 0], m, m) = m I (lib.rs:202-204, 643), each cell's turned blocks
R K_ac R^T times E h (lib.rs:541-543, 650-672), one pair per two nodes of a cell (the
`pattern`, lib.rs:401-422: (p.min(q), p.max(q)) as first met, the block transposed where
p > q). Contacts are left out. The right-hand side is M v + dt M g with v small and random
(the cells at rest: their forces are nothing)."""
import numpy as np

CORNERS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]])
NU = 0.3  # mechanics/src/element.rs:16
DAMPING = 0.3  # mechanics/src/lib.rs:63
GRAVITY = 9.81
WOOD = (600.0, 10e9)  # rho, e (mechanics/src/lib.rs:35)


class System:
    """mechanics::solve::System (mechanics/src/solve.rs:49-59) as arrays: sites (n, 3) int64,
    at (n, 3), diag (n, 3, 3), pairs ij (m, 2) uint32 and their blocks (m, 3, 3)."""

    def __init__(self, sites, at, diag, ij, blocks):
        self.sites = np.ascontiguousarray(sites, np.int64).reshape(-1, 3)
        self.at = np.ascontiguousarray(at, np.float64).reshape(-1, 3)
        self.diag = np.ascontiguousarray(diag, np.float64).reshape(-1, 3, 3)
        self.ij = np.ascontiguousarray(ij, np.uint32).reshape(-1, 2)
        self.blocks = np.ascontiguousarray(blocks, np.float64).reshape(-1, 3, 3)

    @property
    def n(self):
        return len(self.diag)


def unit_k():
    """The unit cube's stiffness (element.rs:49-76): trilinear shape functions, stress of
    strain at Poisson's ratio NU and stiffness 1, 2x2x2 Gauss points; (8, 8, 3, 3), block
    [a][c][i][j]."""
    sign = 2.0 * CORNERS - 1.0
    g = 1.0 / np.sqrt(3.0)
    lam, mu = NU / ((1 + NU) * (1 - 2 * NU)), 1.0 / (2 * (1 + NU))
    d = np.zeros((6, 6))
    d[:3, :3] = lam
    d[np.arange(3), np.arange(3)] += 2 * mu
    d[np.arange(3, 6), np.arange(3, 6)] = mu
    k = np.zeros((24, 24))
    for p in range(8):
        at = sign[p] * g
        b = np.zeros((6, 24))
        for a in range(8):
            s = sign[a]
            grad = np.array([s[kk] / 4.0 * np.prod([1 + s[j] * at[j] for j in range(3) if j != kk]) for kk in range(3)])
            b[0, 3 * a], b[1, 3 * a + 1], b[2, 3 * a + 2] = grad
            b[3, 3 * a + 1], b[3, 3 * a + 2] = grad[2], grad[1]
            b[4, 3 * a], b[4, 3 * a + 2] = grad[2], grad[0]
            b[5, 3 * a], b[5, 3 * a + 1] = grad[1], grad[0]
        k += b.T @ d @ b / 8.0
    return k.reshape(8, 3, 8, 3).transpose(0, 2, 1, 3).copy()


def quat_matrix(q):
    """element.rs:89-95"""
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def system(sites, held_node, h=0.125, mat=WOOD, fill=None, turn=0.0, seed=0, dt=1.0 / 60.0):
    """The system of cells at `sites` (n, 3) int: held_node(corner coordinates (m, 3)) -> bool
    mask of the nodes a held site holds; `turn`: the cells' turns, radians at most (random
    axes); fill: each cell's share of its cube that bears (default 1). Returns (System, b,
    node count, the unknowns' corner coordinates)."""
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
    unit = unit_k()
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
    K = np.einsum("nip,acpq,njq->nacij", R, unit, R, optimize=True)
    kcell = e * fill * h * stiff
    diag = mass[~held][:, None, None] * np.eye(3)[None]
    fa = free[node_of]  # (nc, 8)
    for a in range(8):
        ok = fa[:, a] >= 0
        np.add.at(diag, fa[ok, a], K[ok, a, a] * kcell[ok, None, None])
    # pattern: pairs as first met, cell by cell, a < c; each cell's block added in that order,
    # transposed where the pair's first unknown is the greater
    a_idx, c_idx = np.triu_indices(8, 1)  # (a, c) in the reference's loop order
    p = fa[:, a_idx].reshape(-1)
    q = fa[:, c_idx].reshape(-1)
    ok = (p >= 0) & (q >= 0)
    cell = np.repeat(np.arange(nc), len(a_idx))[ok]
    ac = np.tile(np.arange(len(a_idx)), nc)[ok]
    p, q = p[ok], q[ok]
    lo, hi = np.minimum(p, q), np.maximum(p, q)
    key = lo * (n + 1) + hi
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")  # pairs numbered as first met
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    slot = rank[inv]
    ij = np.stack([lo[first[order]], hi[first[order]]], axis=1)
    m = K[cell, a_idx[ac], c_idx[ac]] * kcell[cell, None, None]
    m = np.where((p > q)[:, None, None], np.transpose(m, (0, 2, 1)), m)
    blocks = np.zeros((len(ij), 3, 3))
    np.add.at(blocks, slot, m)
    s = System(keys[~held], keys[~held] * h, diag, ij.astype(np.uint32).reshape(-1, 2), blocks.reshape(-1, 3, 3))
    v = rng.normal(size=(n, 3)) * 1e-2
    b = (v + np.array([0.0, -GRAVITY * dt, 0.0])) * mass[~held][:, None]
    return s, b, nn, keys[~held]
