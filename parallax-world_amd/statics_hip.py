"""ctypes binding of libstatics_hip.so (include/statics_hip.h).

Used by the tests and bench.py. There is no fallback: if the library is not
built, or no GPU is present, the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libstatics_hip.so")

ST_OK, ST_BAD_INPUT, ST_NO_CONVERGENCE, ST_NOT_SPD, ST_NAN, ST_OOM, ST_DEVICE, ST_NO_STRUCTURE = range(8)
STATUS = {0: "OK", 1: "BAD_INPUT", 2: "NO_CONVERGENCE", 3: "NOT_SPD", 4: "NAN", 5: "OOM", 6: "DEVICE",
          7: "NO_STRUCTURE"}
FP64, MIXED, MIXED32 = 0, 1, 2
COMPLIANCE, SOFTMAX_STRESS = 0, 1
VALUE_ONLY = 256  # or-ed into the objective: J alone

# The entry points include/statics_hip.h declares.
EXPORTS = ["st_create", "st_destroy", "st_last_error", "st_upload", "st_sizes", "st_node_faces", "st_lookup_faces",
           "st_node_held", "st_joined_to_held", "st_set_gone", "st_set_beams", "st_apply", "st_solve",
           "st_set_loads", "st_solve_resident", "st_get_solution", "st_classify", "st_kernel_timing_enable",
           "st_kernel_timing", "st_apply_schedule", "st_bond_forces", "st_bond_used", "st_settle", "st_grid_cells", "st_gather", "st_gather_read", "st_backward", "st_solve_batch", "st_generate", "st_set_stiffness",
           "st_joined_to_held_static", "st_lattice_cells", "st_joined_edges"]
# The entry points include/mechanics_hip.h declares.
ME_EXPORTS = ["me_solver_new", "me_solve", "me_solver_free", "me_last_error", "me_solver_info", "me_kernel_timing"]

UNIT = 0.125  # metres to a level-0 cell (statics/src/lattice.rs:16)


class StLattice(ctypes.Structure):
    _fields_ = [("level", ctypes.c_uint32), ("nnodes", ctypes.c_int64), ("x", ctypes.c_void_p), ("m", ctypes.c_void_p),
                ("node_mat", ctypes.c_void_p), ("held", ctypes.c_void_p), ("around", ctypes.c_void_p),
                ("nbonds", ctypes.c_int64), ("ij", ctypes.c_void_p), ("p", ctypes.c_void_p), ("n", ctypes.c_void_p),
                ("a", ctypes.c_void_p), ("d", ctypes.c_void_p), ("bond_mat", ctypes.c_void_p)]


class GatheredLattice:
    """A lattice as the reference holds it (statics/src/lattice.rs:61-71): node arrays x (n, 3), m,
    node_mat (n, 6: rho, e, g, ft, fc, fs), held, around (n, 2, 3); bond arrays ij (b, 2), p, n (b, 3),
    a, d, bond_mat (b, 6)."""

    def __init__(self, level, nn, nb):
        self.level = level
        self.x, self.m = np.zeros((nn, 3)), np.zeros(nn)
        self.node_mat, self.held = np.zeros((nn, 6)), np.zeros(nn, np.uint8)
        self.around = np.zeros((nn, 2, 3), np.int64)
        self.ij, self.p, self.n = np.zeros((nb, 2), np.int64), np.zeros((nb, 3)), np.zeros((nb, 3))
        self.a, self.d, self.bond_mat = np.zeros(nb), np.zeros(nb), np.zeros((nb, 6))


class StCells(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int64), ("sites", ctypes.c_void_p), ("held", ctypes.c_void_p),
                ("e", ctypes.c_void_p), ("g", ctypes.c_void_p), ("beams", ctypes.c_void_p),
                ("spacing", ctypes.c_double), ("section", ctypes.c_void_p)]


class StGrid(ctypes.Structure):
    _fields_ = [("min", ctypes.c_int32 * 3), ("n", ctypes.c_int32 * 3), ("bricks", ctypes.c_void_p),
                ("nbricks", ctypes.c_int64), ("solid", ctypes.c_void_p), ("nsites", ctypes.c_int64),
                ("solid_bytes", ctypes.c_void_p), ("solid_min", ctypes.c_uint8)]


class StLatticeCells(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int64), ("spacing", ctypes.c_double), ("sites", ctypes.c_void_p),
                ("held", ctypes.c_void_p), ("e", ctypes.c_void_p), ("g", ctypes.c_void_p), ("rho", ctypes.c_void_p),
                ("beams", ctypes.c_void_p), ("strength", ctypes.c_void_p), ("bond_index", ctypes.c_void_p),
                ("section", ctypes.c_void_p)]


class MeSystem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("sites", ctypes.c_void_p), ("at", ctypes.c_void_p), ("diag", ctypes.c_void_p),
                ("npairs", ctypes.c_int64), ("pairs", ctypes.c_void_p), ("blocks", ctypes.c_void_p)]


class StSolveOpts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("tol", ctypes.c_double), ("max_iter", ctypes.c_int32),
                ("refresh", ctypes.c_int32), ("warm_start", ctypes.c_int32), ("release_hierarchy", ctypes.c_int32)]


class StStats(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int32), ("levels", ctypes.c_int32), ("refreshes", ctypes.c_int32),
                ("status", ctypes.c_int32), ("rel_residual", ctypes.c_double), ("setup_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("nodes", ctypes.c_int64), ("dof_nodes", ctypes.c_int64),
                ("cells", ctypes.c_int64), ("bricks", ctypes.c_int64), ("retries", ctypes.c_int32),
                ("first_status", ctypes.c_int32), ("mem_peak", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StGenParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("seed", ctypes.c_uint64),
                ("pillar_spacing", ctypes.c_double), ("pillar_min", ctypes.c_int32), ("pillar_max", ctypes.c_int32),
                ("slab_min", ctypes.c_int32), ("slab_max", ctypes.c_int32), ("hole_p", ctypes.c_double),
                ("sigma", ctypes.c_double)]


_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        sig = {
            "st_create": [ctypes.POINTER(vp), ctypes.c_int],
            "st_destroy": [vp],
            "st_upload": [vp, ctypes.POINTER(StCells)],
            "st_sizes": [vp, vp, vp, vp],
            "st_node_faces": [vp, vp],
            "st_lookup_faces": [vp, i64, vp, vp],
            "st_node_held": [vp, vp],
            "st_joined_to_held": [vp, vp, vp, ctypes.c_int],
            "st_set_gone": [vp, vp],
            "st_set_beams": [vp, vp],
            "st_set_stiffness": [vp, vp, vp],
            "st_apply": [vp, vp, vp],
            "st_solve": [vp, vp, ctypes.POINTER(StSolveOpts), vp, ctypes.POINTER(StStats)],
            "st_bond_forces": [vp, vp, vp],
            "st_bond_used": [vp, vp, ctypes.c_int, vp, vp],
            "st_lattice_cells": [ctypes.POINTER(StLattice), ctypes.POINTER(StLatticeCells), ctypes.c_char_p,
                                 ctypes.c_int64],
            "st_grid_cells": [vp, ctypes.POINTER(StGrid), ctypes.c_int32, dbl, vp, vp, vp, vp],
            "st_gather": [vp, ctypes.POINTER(StLattice), ctypes.c_int, vp, vp],
            "st_gather_read": [vp] * 13,
            "st_settle": [vp, vp, ctypes.c_int, vp, ctypes.POINTER(StSolveOpts), vp, vp, vp, ctypes.POINTER(StStats)],
            "st_backward": [vp, ctypes.c_int, dbl, ctypes.POINTER(StSolveOpts), vp, vp, ctypes.POINTER(StStats)],
            "st_solve_batch": [vp, i32, vp, vp, ctypes.POINTER(StSolveOpts), vp, vp],
            "st_generate": [ctypes.POINTER(StGenParams), vp, vp],
            "st_set_loads": [vp, vp],
            "st_solve_resident": [vp, ctypes.POINTER(StSolveOpts), ctypes.POINTER(StStats)],
            "st_get_solution": [vp, vp],
            "st_classify": [vp, vp],
            "st_kernel_timing_enable": [vp, ctypes.c_int, ctypes.c_int],
            "st_kernel_timing": [vp, vp, vp, vp],
            "st_apply_schedule": [vp, ctypes.c_int, vp, vp, vp, vp],
            "st_joined_to_held_static": [vp, vp, vp, vp],
            "st_joined_edges": [vp, ctypes.c_int64, vp, ctypes.c_int64, vp, vp, vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.st_last_error.argtypes = [vp]
        L.st_last_error.restype = ctypes.c_char_p
        me = {
            "me_solver_new": [ctypes.POINTER(vp), ctypes.POINTER(MeSystem), dbl, ctypes.c_int],
            "me_solve": [vp, ctypes.POINTER(MeSystem), vp, vp, dbl, vp],
            "me_solver_free": [vp],
            "me_solver_info": [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp],
            "me_kernel_timing": [vp, ctypes.c_int, vp, vp, vp],
        }
        for name, args in me.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.me_last_error.argtypes = [vp]
        L.me_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class StaticsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def generate(nx, ny, nz, seed=0, pillar_spacing=8.0, pillar_min=2, pillar_max=6, slab_min=8, slab_max=16,
             hole_p=0.2, sigma=0.0):
    """The seeded pillar/slab structure: occ[x, y, z] in {0 empty, 1 free, 2 held} and theta (log-stiffness)."""
    occ = np.zeros((nx, ny, nz), np.uint8)
    theta = np.zeros((nx, ny, nz), np.float32)
    p = StGenParams(nx, ny, nz, seed, pillar_spacing, pillar_min, pillar_max, slab_min, slab_max, hole_p, sigma)
    rc = lib().st_generate(ctypes.byref(p), _p(occ), _p(theta))
    if rc != 0:
        raise StaticsError(rc, "st_generate")
    return occ, theta


def cells_from_occ(occ, theta=None, e0=1.0, g_over_e=0.4):
    """Solid cells of occ in the reference's `solids` order (x outer, z inner): sites, held, e, g."""
    sites = np.argwhere(occ > 0).astype(np.int32)
    held = (occ[occ > 0] == 2).astype(np.uint8)
    th = np.zeros(len(sites)) if theta is None else theta[occ > 0].astype(np.float64)
    e = e0 * np.exp(th)
    g = g_over_e * e
    return sites, held, e, g


def gravity_loads(faces, held_or_gone, count_cells, weight=1.0, axis=2):
    """Loads of weight w per free cell, a sixth on each of its faces, along -axis."""
    f = np.zeros((len(faces), 6))
    f[:, axis] = -weight * count_cells / 6.0
    f[held_or_gone.astype(bool)] = 0.0
    return f


def _strength(strength, ncells):
    a = np.ascontiguousarray(strength, np.float64)
    if a.size == ncells * 45:
        return a.reshape(ncells, 45), 1
    return a.reshape(ncells, 3), 0


def lattice_arrays(lat):
    """A lattice's arrays as StLattice wants them, and the StLattice (which points into them)."""
    arr = {k: np.ascontiguousarray(getattr(lat, k), t) for k, t in
           [("x", np.float64), ("m", np.float64), ("node_mat", np.float64), ("held", np.uint8),
            ("around", np.int64), ("ij", np.int64), ("p", np.float64), ("n", np.float64), ("a", np.float64),
            ("d", np.float64), ("bond_mat", np.float64)]}
    st = StLattice(int(lat.level), len(arr["m"]), _p(arr["x"]), _p(arr["m"]), _p(arr["node_mat"]), _p(arr["held"]),
                   _p(arr["around"]), len(arr["a"]), _p(arr["ij"]), _p(arr["p"]), _p(arr["n"]), _p(arr["a"]),
                   _p(arr["d"]), _p(arr["bond_mat"]))
    return arr, st


def lattice_cells(lat, sections=True):
    """st_lattice_cells: the cells a built or gathered lattice comes from, as a dict of arrays
    (sites, held, e, g, rho, beams, strength (cells, 15, 3), section (cells, 15), bond_index) and
    spacing. Raises StaticsError with the library's reason for a lattice neither build nor
    gather::measured could have made. sections=False asks for no sections: a lattice whose
    beams have sections of their own is then refused."""
    arr, st = lattice_arrays(lat)
    out = StLatticeCells()
    msg = ctypes.create_string_buffer(512)
    rc = lib().st_lattice_cells(ctypes.byref(st), ctypes.byref(out), msg, 512)
    if rc != 0:
        raise StaticsError(rc, msg.value.decode())
    n = out.n_cells
    res = {"sites": np.zeros((n, 3), np.int32), "held": np.zeros(n, np.uint8), "e": np.zeros(n), "g": np.zeros(n),
           "rho": np.zeros(n), "beams": np.zeros(n, np.uint16), "strength": np.zeros((n, 15, 3)),
           "section": np.zeros((n, 15)), "bond_index": np.zeros(len(arr["a"]), np.int64)}
    for k in ("sites", "held", "e", "g", "rho", "beams", "strength", "section", "bond_index"):
        setattr(out, k, res[k].ctypes.data)
    if not sections:
        out.section = None
        del res["section"]
    rc = lib().st_lattice_cells(ctypes.byref(st), ctypes.byref(out), msg, 512)
    if rc != 0:
        raise StaticsError(rc, msg.value.decode())
    res["spacing"] = out.spacing
    return res


class Context:
    """One device context (st_create .. st_destroy)."""

    def __init__(self, device=0):
        self.L = lib()
        h = ctypes.c_void_p()
        rc = self.L.st_create(ctypes.byref(h), device)
        if rc != 0:
            raise StaticsError(rc, "st_create (is a GPU visible?)")
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            self.L.st_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise StaticsError(rc, f"{what}: {self.L.st_last_error(self.h).decode()}")

    @staticmethod
    def make_cells(sites, held, e, g, beams=None, spacing=UNIT, section=None):
        keep = [np.ascontiguousarray(sites, np.int32), np.ascontiguousarray(held, np.uint8),
                np.ascontiguousarray(e, np.float64), np.ascontiguousarray(g, np.float64),
                None if beams is None else np.ascontiguousarray(beams, np.uint16),
                None if section is None else np.ascontiguousarray(section, np.float64).reshape(-1, 15)]
        if keep[5] is not None and len(keep[5]) != len(keep[0]):
            raise ValueError("section: 15 per cell")
        c = StCells(len(keep[0]), keep[0].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data, keep[3].ctypes.data,
                    None if keep[4] is None else keep[4].ctypes.data, spacing,
                    None if keep[5] is None else keep[5].ctypes.data)
        return c, keep

    def upload(self, sites, held, e, g, beams=None, spacing=UNIT, section=None):
        """st_upload. section: (cells, 15) beam sections over SECTION s^2, or None (all 1)."""
        c, keep = self.make_cells(sites, held, e, g, beams, spacing, section)
        self._check(self.L.st_upload(self.h, ctypes.byref(c)), "st_upload")
        self._keep = keep
        n, nc, nb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._check(self.L.st_sizes(self.h, ctypes.byref(n), ctypes.byref(nc), ctypes.byref(nb)), "st_sizes")
        self.nodes, self.cells, self.bricks = n.value, nc.value, nb.value
        return self

    def node_faces(self):
        f = np.zeros((self.nodes, 4), np.int32)
        self._check(self.L.st_node_faces(self.h, _p(f)), "st_node_faces")
        return f

    def lookup_faces(self, faces):
        faces = np.ascontiguousarray(faces, np.int32)
        idx = np.zeros(len(faces), np.int64)
        self._check(self.L.st_lookup_faces(self.h, len(faces), _p(faces), _p(idx)), "st_lookup_faces")
        return idx

    def node_held(self):
        h = np.zeros(self.nodes, np.uint8)
        self._check(self.L.st_node_held(self.h, _p(h)), "st_node_held")
        return h

    def joined_to_held(self, beams=None, keep_gone=False):
        j = np.zeros(self.nodes, np.uint8)
        b = None if beams is None else np.ascontiguousarray(beams, np.uint16)
        self._check(self.L.st_joined_to_held(self.h, _p(b), _p(j), 1 if keep_gone else 0), "st_joined_to_held")
        return j

    def joined_edges(self, held, ij, alive=None):
        """bodies::joined_to_held(nodes, bonds, alive) over a lattice's own nodes and bonds."""
        held = np.ascontiguousarray(held, np.uint8)
        ij = np.ascontiguousarray(ij, np.int64).reshape(-1, 2)
        a = None if alive is None else np.ascontiguousarray(alive, np.uint8)
        out = np.zeros(len(held), np.uint8)
        self._check(self.L.st_joined_edges(self.h, len(held), _p(held), len(ij), _p(ij), _p(a), _p(out)),
                    "st_joined_edges")
        return out

    def set_gone(self, gone):
        g = None if gone is None else np.ascontiguousarray(gone, np.uint8)
        self._check(self.L.st_set_gone(self.h, _p(g)), "st_set_gone")

    def set_beams(self, beams):
        b = None if beams is None else np.ascontiguousarray(beams, np.uint16)
        self._check(self.L.st_set_beams(self.h, _p(b)), "st_set_beams")

    def apply(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(self.nodes, 6)
        y = np.zeros_like(x)
        self._check(self.L.st_apply(self.h, _p(x), _p(y)), "st_apply")
        return y

    def solve(self, loads, precision=FP64, tol=1e-8, max_iter=10000, refresh=0, x0=None, check=True,
              release_hierarchy=False):
        loads = np.ascontiguousarray(loads, np.float64).reshape(self.nodes, 6)
        x = np.zeros_like(loads) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        o = StSolveOpts(precision, tol, max_iter, refresh, 0 if x0 is None else 1, 1 if release_hierarchy else 0)
        s = StStats()
        rc = self.L.st_solve(self.h, _p(loads), ctypes.byref(o), _p(x), ctypes.byref(s))
        if check:
            self._check(rc, "st_solve")
        if rc != 0 and s.status == 0:
            s.status = rc
        return x, s

    def bond_forces(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(self.nodes, 6)
        out = np.zeros((len(self._keep[0]), 15, 4, 3))
        self._check(self.L.st_bond_forces(self.h, _p(x), _p(out)), "st_bond_forces")
        return out

    def grid_cells(self, min_site, n, bricks, solid, nsites, level=0, hold_radius=0.0, solid_min=None):
        """solids + hold_bulk for a lattice::Grid (solid: its bits as uint32 words) or, with solid_min
        set, a VoxelGrid (solid: a byte per site): (sites (m, 3) int32, held (m,), index (m,))."""
        bricks = np.ascontiguousarray(bricks, np.int32)
        if solid_min is None:
            solid = np.ascontiguousarray(solid, np.uint32)
            gr = StGrid((ctypes.c_int32 * 3)(*min_site), (ctypes.c_int32 * 3)(*n), _p(bricks), len(bricks),
                        _p(solid), int(nsites), None, 0)
        else:
            solid = np.ascontiguousarray(solid, np.uint8)
            gr = StGrid((ctypes.c_int32 * 3)(*min_site), (ctypes.c_int32 * 3)(*n), _p(bricks), len(bricks), None,
                        int(nsites), _p(solid), int(solid_min))
        m = ctypes.c_int64()
        self._check(self.L.st_grid_cells(self.h, ctypes.byref(gr), level, hold_radius, ctypes.byref(m), None, None,
                                         None), "st_grid_cells")
        sites = np.zeros((m.value, 3), np.int32)
        held = np.zeros(m.value, np.uint8)
        index = np.zeros(m.value, np.int64)
        self._check(self.L.st_grid_cells(self.h, ctypes.byref(gr), level, hold_radius, ctypes.byref(m), _p(sites),
                                         _p(held), _p(index)), "st_grid_cells")
        return sites, held, index

    def gather(self, fine, against_full=True):
        """gather::gathered (against_full) or measured(fine, false) (statics/src/gather.rs:193-395) of a
        lattice with the attributes GatheredLattice names: the next level up, as a GatheredLattice."""
        arr, lat = lattice_arrays(fine)
        nn, nb = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.L.st_gather(self.h, ctypes.byref(lat), 1 if against_full else 0, ctypes.byref(nn),
                                     ctypes.byref(nb)), "st_gather")
        out = GatheredLattice(0, nn.value, nb.value)
        level = ctypes.c_uint32()
        self._check(self.L.st_gather_read(self.h, ctypes.byref(level), _p(out.x), _p(out.m), _p(out.node_mat),
                                          _p(out.held), _p(out.around), _p(out.ij), _p(out.p), _p(out.n),
                                          _p(out.a), _p(out.d), _p(out.bond_mat)), "st_gather_read")
        out.level = level.value
        return out

    def bond_used(self, strength, x):
        """Per uploaded cell [15][4] used shares; strength: (cells, 3) ft, fc, fs per cell, or
        (cells, 15, 3) per beam."""
        x = np.ascontiguousarray(x, np.float64).reshape(self.nodes, 6)
        s, per_beam = _strength(strength, len(self._keep[0]))
        out = np.zeros((len(self._keep[0]), 15, 4))
        self._check(self.L.st_bond_used(self.h, _p(s), per_beam, _p(x), _p(out)), "st_bond_used")
        return out

    def settle(self, strength, loads, precision=FP64, tol=1e-6, max_iter=10000, warm_start=False, refresh=0):
        """bodies::settle on the device: (gone per node, final beam masks per cell, rounds, stats).
        tol: the reference settles at 1e-6 (statics/src/bodies.rs:432); refresh -1: its stopping rule."""
        s, per_beam = _strength(strength, len(self._keep[0]))
        loads = np.ascontiguousarray(loads, np.float64).reshape(self.nodes, 6)
        o = StSolveOpts(precision, tol, max_iter, refresh, 1 if warm_start else 0, 0)
        gone = np.zeros(self.nodes, np.uint8)
        beams = np.zeros(len(self._keep[0]), np.uint16)
        rounds = ctypes.c_int32()
        st = StStats()
        self._check(self.L.st_settle(self.h, _p(s), per_beam, _p(loads), ctypes.byref(o), _p(gone), _p(beams),
                                     ctypes.byref(rounds), ctypes.byref(st)), "st_settle")
        return gone, beams, rounds.value, st

    def backward(self, objective=COMPLIANCE, beta=20.0, precision=FP64, tol=1e-8, max_iter=10000,
                 release_hierarchy=False):
        o = StSolveOpts(precision, tol, max_iter, 0, 0, 1 if release_hierarchy else 0)
        J = ctypes.c_double()
        d = np.zeros(len(self._keep[0]))
        s = StStats()
        self._check(self.L.st_backward(self.h, objective, beta, ctypes.byref(o), ctypes.byref(J), _p(d),
                                       ctypes.byref(s)), "st_backward")
        return J.value, d, s

    def objective(self, objective=COMPLIANCE, beta=20.0):
        """J of the last solve alone (no adjoint solve)."""
        o = StSolveOpts(FP64, 1e-8, 1, 0, 0, 0)
        J = ctypes.c_double()
        self._check(self.L.st_backward(self.h, objective | VALUE_ONLY, beta, ctypes.byref(o), ctypes.byref(J),
                                       None, None), "st_backward")
        return J.value

    def set_stiffness(self, e, g):
        """New e, g per uploaded cell (upload order), keeping the lattice."""
        e = np.ascontiguousarray(e, np.float64)
        g = np.ascontiguousarray(g, np.float64)
        self._check(self.L.st_set_stiffness(self.h, _p(e), _p(g)), "st_set_stiffness")

    def apply_schedule(self, fp64=True):
        """(index, kernel, description, dispatches per application) of the fine operator's tuned schedule."""
        i, k, d, n = ctypes.c_int(), ctypes.c_char_p(), ctypes.c_char_p(), ctypes.c_int()
        self._check(self.L.st_apply_schedule(self.h, 1 if fp64 else 0, ctypes.byref(i), ctypes.byref(k),
                                             ctypes.byref(d), ctypes.byref(n)), "st_apply_schedule")
        return i.value, k.value.decode(), d.value.decode(), n.value

    def solve_batch(self, cell_sets, loads=None, precision=MIXED, tol=1e-8, max_iter=10000, want_x=True):
        """cell_sets: list of (sites, held, e, g). Gravity (weight 1 per free cell, -z) where loads[i] is None."""
        n = len(cell_sets)
        structs, keeps = [], []
        for cs in cell_sets:
            c, k = self.make_cells(*cs)
            structs.append(c)
            keeps.append(k)
        arr = (ctypes.POINTER(StCells) * n)(*[ctypes.pointer(c) for c in structs])
        lptr = (ctypes.c_void_p * n)()
        lkeep = []
        if loads is not None:
            for i, l in enumerate(loads):
                if l is not None:
                    a = np.ascontiguousarray(l, np.float64)
                    lkeep.append(a)
                    lptr[i] = a.ctypes.data
        stats = (StStats * n)()
        xs = [None] * n
        xptr = (ctypes.c_void_p * n)()
        o = StSolveOpts(precision, tol, max_iter, 0, 0, 0)
        # node counts are known only after the upload: a first call without outputs would cost a solve,
        # so x comes back through a second, zero-iteration-free route: size by an upper bound of 6 faces per cell
        if want_x:
            for i, k in enumerate(keeps):
                xs[i] = np.zeros((6 * int((k[1] == 0).sum()) + 1, 6))
                xptr[i] = xs[i].ctypes.data
        rc = self.L.st_solve_batch(self.h, n, arr, lptr, ctypes.byref(o), xptr if want_x else None, stats)
        self._check(rc, "st_solve_batch")
        if want_x:
            xs = [x[:stats[i].nodes] for i, x in enumerate(xs)]
        return xs, [stats[i] for i in range(n)]


class MechSolver:
    """mechanics::solve::Solver on the device (include/mechanics_hip.h): made for a system
    (sites, at, diag, pairs ij, blocks as numpy arrays, the shapes of mechanics::solve::System),
    then solve(system, b, x0, tol) -> (x, rounds or None)."""

    @staticmethod
    def _system(s):
        keep = [np.ascontiguousarray(s.sites, np.int64), np.ascontiguousarray(s.at, np.float64),
                np.ascontiguousarray(s.diag, np.float64), np.ascontiguousarray(s.ij, np.uint32),
                np.ascontiguousarray(s.blocks, np.float64)]
        m = MeSystem(len(keep[2]), keep[0].ctypes.data, keep[1].ctypes.data, keep[2].ctypes.data, len(keep[3]),
                     keep[3].ctypes.data if len(keep[3]) else None, keep[4].ctypes.data if len(keep[4]) else None)
        return m, keep

    def __init__(self, system, size, device=0):
        self.L = lib()
        m, keep = self._system(system)
        h = ctypes.c_void_p()
        rc = self.L.me_solver_new(ctypes.byref(h), ctypes.byref(m), size, device)
        if rc != 0:
            raise StaticsError(rc, "me_solver_new (the reason is on stderr)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.me_solver_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, system, b, x0=None, tol=1e-3):
        m, keep = self._system(system)
        b = np.ascontiguousarray(b, np.float64).reshape(-1, 3)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).reshape(-1, 3).copy()
        rounds = ctypes.c_int32()
        rc = self.L.me_solve(self.h, ctypes.byref(m), _p(b), _p(x), tol, ctypes.byref(rounds))
        if rc != 0:
            raise StaticsError(rc, f"me_solve: {self.L.me_last_error(self.h).decode()}")
        return x, (None if rounds.value < 0 else rounds.value)

    def kernel_timing(self, enable):
        """me_kernel_timing: (launches, total ms, algorithmic bytes per launch) of the CG's K p."""
        n, ms, b = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        rc = self.L.me_kernel_timing(self.h, enable, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b))
        if rc != 0:
            raise StaticsError(rc, "me_kernel_timing")
        return n.value, ms.value, b.value

    def info(self, cap=64):
        count = ctypes.c_int32()
        parts, pairs = np.zeros(cap, np.int64), np.zeros(cap, np.int64)
        steps, direct = np.zeros(cap), np.zeros(cap, np.int32)
        step0, setup_ms, solve_ms = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        rc = self.L.me_solver_info(self.h, cap, ctypes.byref(count), _p(parts), _p(pairs), _p(steps), _p(direct),
                                   ctypes.byref(step0), ctypes.byref(setup_ms), ctypes.byref(solve_ms))
        if rc != 0:
            raise StaticsError(rc, "me_solver_info")
        n = count.value
        return {"levels": [(int(parts[l]), int(pairs[l]), float(steps[l]), bool(direct[l])) for l in range(n)],
                "step0": step0.value, "setup_ms": setup_ms.value, "solve_ms": solve_ms.value}
