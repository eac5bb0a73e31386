"""The oracle: ctypes binding of the CPU FP64 restatement in statics_oracle.cpp.

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker. The product never loads it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "statics_oracle.cpp")
MECH_SRC = os.path.join(HERE, "mechanics_oracle.cpp")
LIB_PATH = os.path.join(HERE, "liboracle.so")
CPU_TAG = os.path.join(HERE, "liboracle.cpu")  # the host CPU the library was built for (-march=native)

# Materials (statics/src/model.rs:53-58): rho, e, g, ft, fc, fs
STONE = (2500.0, 30e9, 12e9, 5e6, 100e6, 10e6)
WOOD = (600.0, 10e9, 4e9, 40e6, 30e6, 6e6)
MASONRY = (2000.0, 5e9, 2e9, 0.2e6, 10e6, 0.4e6)
SOIL = (1800.0, 50e6, 20e6, 5e3, 0.3e6, 20e3)
STEEL = (7850.0, 200e9, 80e9, 250e6, 250e6, 145e6)
LEAVES = (60.0, 0.2e9, 0.08e9, 5e6, 5e6, 1e6)
GRAVITY = 9.81
UNIT = 0.125


def host_cpu():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def build():
    """Compiles the restatement with the host's g++ (-O3 -march=native, as timed; no FMA
    contraction, as Rust does not contract)."""
    tmp = LIB_PATH + ".tmp"
    cmd = ["g++", "-O3", "-march=native", "-ffp-contract=off", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o",
           tmp, SRC, MECH_SRC]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    with open(CPU_TAG, "w") as f:
        f.write(host_cpu())
    return LIB_PATH


def _stale():
    if not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(os.path.getmtime(SRC),
                                                                       os.path.getmtime(MECH_SRC)):
        return True
    try:
        with open(CPU_TAG) as f:
            return f.read().strip() != host_cpu()
    except OSError:
        return True


_lib = None


def lib():
    global _lib
    if _lib is None:
        if _stale():
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.orc_build.restype = vp
        L.orc_build.argtypes = [ctypes.c_uint32, ctypes.c_size_t, vp, vp, vp]
        L.orc_free.argtypes = [vp]
        L.orc_hold_bulk.argtypes = [ctypes.c_uint32, ctypes.c_size_t, vp, ctypes.c_double, vp]
        L.orc_sizes.argtypes = [vp, vp, vp]
        L.orc_nodes.argtypes = [vp, vp, vp, vp, vp]
        L.orc_bonds.argtypes = [vp] * 9
        L.orc_joined_to_held.argtypes = [vp, vp, vp]
        L.orc_motions.restype = ctypes.c_int
        L.orc_motions.argtypes = [vp, vp, vp, vp, ctypes.c_double, ctypes.c_int, vp, vp, vp, vp, vp, vp]
        L.orc_apply.argtypes = [vp, vp, vp, vp, vp]
        L.orc_moved.argtypes = [vp, vp, vp]
        L.orc_forces.argtypes = [vp, vp, vp, ctypes.c_double, vp]
        L.orc_used.argtypes = [vp, vp, vp, vp]
        L.orc_used_standing.argtypes = [vp, ctypes.c_double, vp]
        L.orc_settle.argtypes = [vp, ctypes.c_double, vp, vp, vp, vp]
        L.orc_scale_cell.argtypes = [vp, ctypes.c_int64, ctypes.c_double]
        L.orc_as_generated.argtypes = [vp, vp, ctypes.c_double]
        L.orc_invert6.argtypes = [vp, vp]
        L.orc_coarse_site.argtypes = [vp, ctypes.c_uint32, vp]
        L.orc_spacing.restype = ctypes.c_double
        L.orc_spacing.argtypes = [ctypes.c_uint32]
        L.orc_gathered.restype = vp
        L.orc_gathered.argtypes = [vp, ctypes.c_int]
        L.orc_level.restype = ctypes.c_uint32
        L.orc_level.argtypes = [vp]
        L.orc_mats.argtypes = [vp, vp, vp]
        L.orc_gather_solve.argtypes = [ctypes.c_size_t, ctypes.c_size_t, vp, vp]
        L.orc_beams.restype = ctypes.c_size_t
        L.orc_beams.argtypes = [ctypes.c_size_t, vp, vp, vp, vp]
        L.orc_cell_stiffness.restype = ctypes.c_size_t
        L.orc_cell_stiffness.argtypes = [vp, ctypes.c_size_t, vp, vp]
        L.orc_energy.argtypes = [vp, vp, vp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_progress.argtypes = [ctypes.c_int]
        L.orc_threads.restype = ctypes.c_int
        L.mor_new.restype = vp
        L.mor_new.argtypes = [ctypes.c_int64, vp, vp, vp, ctypes.c_int64, vp, vp, ctypes.c_double]
        L.mor_free.argtypes = [vp]
        L.mor_solve.restype = ctypes.c_int64
        L.mor_solve.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int64, vp, vp, vp, vp, ctypes.c_double]
        L.mor_precondition.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int64, vp, vp, vp, vp]
        L.mor_times.argtypes = [ctypes.c_int64, vp, ctypes.c_int64, vp, vp, vp, vp]
        L.mor_levels.restype = ctypes.c_int
        L.mor_levels.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp]
        L.mor_level_map.argtypes = [vp, ctypes.c_int, vp, vp]
        L.mor_lift.argtypes = [vp, vp, vp, vp]
        L.mor_inverse6.argtypes = [vp, vp]
        L.mor_unit_k.argtypes = [vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def hold_bulk(level, sites, radius):
    sites = np.ascontiguousarray(sites, np.int64)
    held = np.zeros(len(sites), np.uint8)
    lib().orc_hold_bulk(level, len(sites), _p(sites), radius, _p(held))
    return held


class Lattice:
    """The reference's Lattice (statics/src/lattice.rs:61-71), built by the restatement of `build`."""

    def __init__(self, level, sites, mats, held):
        self.level = level
        self.sites = np.ascontiguousarray(sites, np.int64)
        self.mats = np.ascontiguousarray(np.broadcast_to(np.asarray(mats, np.float64), (len(self.sites), 6)))
        self.held_cells = np.ascontiguousarray(held, np.uint8)
        self.h = ctypes.c_void_p(lib().orc_build(level, len(self.sites), _p(self.sites), _p(self.mats),
                                                 _p(self.held_cells)))
        self._read()

    @classmethod
    def _of_handle(cls, h):
        lat = cls.__new__(cls)
        lat.h = h
        lat.level = lib().orc_level(h)
        lat.sites = lat.mats = lat.held_cells = None
        lat._read()
        return lat

    def gathered(self, against_full=True):
        """gather::gathered (statics/src/gather.rs:193-195), or measured(fine, false) (:228)."""
        return Lattice._of_handle(ctypes.c_void_p(lib().orc_gathered(self.h, 1 if against_full else 0)))

    def _read(self):
        nn, nb = ctypes.c_size_t(), ctypes.c_size_t()
        lib().orc_sizes(self.h, ctypes.byref(nn), ctypes.byref(nb))
        self.n_nodes, self.n_bonds = nn.value, nb.value
        self.x = np.zeros((self.n_nodes, 3))
        self.m = np.zeros(self.n_nodes)
        self.held = np.zeros(self.n_nodes, np.uint8)
        self.around = np.zeros((self.n_nodes, 2, 3), np.int64)
        lib().orc_nodes(self.h, _p(self.x), _p(self.m), _p(self.held), _p(self.around))
        nbm = self.n_bonds
        self.ij = np.zeros((nbm, 2), np.int64)
        self.p = np.zeros((nbm, 3))
        self.n = np.zeros((nbm, 3))
        self.a = np.zeros(nbm)
        self.d = np.zeros(nbm)
        self.e = np.zeros(nbm)
        self.g = np.zeros(nbm)
        self.beam_cell = np.zeros(nbm // 4, np.int64)
        lib().orc_bonds(self.h, _p(self.ij), _p(self.p), _p(self.n), _p(self.a), _p(self.d), _p(self.e), _p(self.g),
                        _p(self.beam_cell))
        self.node_mat = np.zeros((self.n_nodes, 6))
        self.bond_mat = np.zeros((nbm, 6))
        lib().orc_mats(self.h, _p(self.node_mat), _p(self.bond_mat))

    def __del__(self):
        try:
            lib().orc_free(self.h)
        except Exception:
            pass

    def _alive_gone(self, alive, gone):
        alive = np.ones(self.n_bonds, np.uint8) if alive is None else np.ascontiguousarray(alive, np.uint8)
        gone = np.zeros(self.n_nodes, np.uint8) if gone is None else np.ascontiguousarray(gone, np.uint8)
        return alive, gone

    def joined_to_held(self, alive=None):
        alive, _ = self._alive_gone(alive, None)
        out = np.zeros(self.n_nodes, np.uint8)
        lib().orc_joined_to_held(self.h, _p(alive), _p(out))
        return out

    def apply(self, x, alive=None, gone=None):
        alive, gone = self._alive_gone(alive, gone)
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        lib().orc_apply(self.h, _p(alive), _p(gone), _p(x), _p(y))
        return y

    def motions(self, loads, alive=None, gone=None, tol=1e-6, max_iter=100000):
        """bodies::motions with the tolerance as a parameter. Returns x (n_nodes x 6) and stats."""
        alive, gone = self._alive_gone(alive, gone)
        loads = np.ascontiguousarray(loads, np.float64)
        x = np.zeros((self.n_nodes, 6))
        it, lv = ctypes.c_int(), ctypes.c_int()
        work, ts, tv = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        rc = lib().orc_motions(self.h, _p(alive), _p(gone), _p(loads), tol, max_iter, _p(x), ctypes.byref(it),
                               ctypes.byref(work), ctypes.byref(ts), ctypes.byref(tv), ctypes.byref(lv))
        return x, dict(iters=it.value, work=work.value, setup_s=ts.value, solve_s=tv.value, levels=lv.value,
                       capped=rc != 0)

    def moved(self, x):
        f = np.zeros((self.n_bonds, 3))
        lib().orc_moved(self.h, _p(np.ascontiguousarray(x, np.float64)), _p(f))
        return f

    def forces(self, alive=None, gone=None, tol=1e-6):
        alive, gone = self._alive_gone(alive, gone)
        f = np.zeros((self.n_bonds, 3))
        lib().orc_forces(self.h, _p(alive), _p(gone), tol, _p(f))
        return f

    def used(self, f, strength=None):
        u = np.zeros(self.n_bonds)
        s = None if strength is None else np.ascontiguousarray(strength, np.float64)
        lib().orc_used(self.h, _p(np.ascontiguousarray(f, np.float64)), _p(s), _p(u))
        return u

    def used_standing(self, tol=1e-6):
        u = np.zeros(self.n_bonds)
        lib().orc_used_standing(self.h, tol, _p(u))
        return u

    def settle(self, tol=1e-6):
        gone = np.zeros(self.n_nodes, np.uint8)
        fallen, rounds, work = ctypes.c_double(), ctypes.c_int(), ctypes.c_double()
        lib().orc_settle(self.h, tol, _p(gone), ctypes.byref(fallen), ctypes.byref(rounds), ctypes.byref(work))
        return dict(fallen=fallen.value, gone=gone, rounds=rounds.value, work=work.value)

    def cell_stiffness(self, cap=4096):
        """Each measured cell's stiffness between its handles (statics/src/gather.rs:291-303), as
        measured(self, false) finds it: a list of (3 nh, 3 nh) arrays."""
        out = np.zeros(cap * (1 + 18 * 18))
        n = ctypes.c_size_t()
        cells = lib().orc_cell_stiffness(self.h, cap, _p(out), ctypes.byref(n))
        ks, at = [], 0
        for _ in range(cells):
            m = 3 * int(out[at])
            ks.append(out[at + 1:at + 1 + m * m].reshape(m, m).copy())
            at += 1 + m * m
        return ks

    def energy(self, f):
        """Bond::energy (statics/src/model.rs:132-136) of each bond bearing f (bonds, 3)."""
        out = np.zeros(self.n_bonds)
        lib().orc_energy(self.h, _p(np.ascontiguousarray(f, np.float64)), _p(out))
        return out

    def as_generated(self, generated, tol=1e-6):
        """bodies::as_generated (statics/src/bodies.rs:604-625): each beam's bonds made MARGIN times
        as strong as generated's same beam needs standing (tol: used_standing's solve)."""
        lib().orc_as_generated(self.h, generated.h, tol)
        self._read()

    def scale_cell(self, cell, by):
        lib().orc_scale_cell(self.h, int(cell), float(by))

    def gravity(self, axis=1, g=GRAVITY):
        """forces' loads (statics/src/bodies.rs:408): -m g along axis (the reference: y)."""
        f = np.zeros((self.n_nodes, 6))
        f[:, axis] = -self.m * g
        return f


def invert6(m):
    m = np.ascontiguousarray(m, np.float64)
    out = np.zeros((6, 6))
    lib().orc_invert6(_p(m), _p(out))
    return out


def gather_solve(a, rhs):
    """gather.rs:165-186 on a (n, n + rhs) matrix: the (n, rhs) solutions."""
    a = np.ascontiguousarray(a, np.float64)
    n = a.shape[0]
    x = np.zeros((n, rhs))
    lib().orc_gather_solve(n, rhs, _p(a), _p(x))
    return x


def beams(at, k, joined):
    """gather.rs:116-161: [(a, b, stiffness)] for handles at `at`, stiffness k, joined (nh, nh)."""
    at = np.ascontiguousarray(at, np.float64)
    nh = len(at)
    out = np.zeros((max(nh * (nh - 1) // 2, 1), 3))
    m = lib().orc_beams(nh, _p(at), _p(np.ascontiguousarray(k, np.float64)),
                        _p(np.ascontiguousarray(joined, np.uint8)), _p(out))
    return [(int(a), int(b), float(v)) for a, b, v in out[:m]]


def coarse_site(p, l):
    p = np.ascontiguousarray(p, np.int64)
    out = np.zeros(3, np.int64)
    lib().orc_coarse_site(_p(p), l, _p(out))
    return out


def spacing(level):
    """statics/src/lattice.rs:34-36"""
    return lib().orc_spacing(level)


# ---------------------------------------------------------------- mechanics::solve
class MechSystem:
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

    def times(self, x):
        """System::times (solve.rs:61-72)."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        y = np.zeros_like(x)
        lib().mor_times(self.n, _p(self.diag), len(self.ij), _p(self.ij), _p(self.blocks), _p(x), _p(y))
        return y

    def dense(self):
        """The system as a dense (3n, 3n) matrix (tests' brute force)."""
        n = self.n
        a = np.zeros((3 * n, 3 * n))
        for i in range(n):
            a[3 * i:3 * i + 3, 3 * i:3 * i + 3] += self.diag[i]
        for (i, j), b in zip(self.ij, self.blocks):
            a[3 * i:3 * i + 3, 3 * j:3 * j + 3] += b
            a[3 * j:3 * j + 3, 3 * i:3 * i + 3] += b.T
        return a


class MechSolver:
    """mechanics::solve::Solver (solve.rs:379-537): new(system, size), solve(system, b, x, tol)."""

    def __init__(self, system, size):
        s = system
        self.h = ctypes.c_void_p(lib().mor_new(s.n, _p(s.sites), _p(s.at), _p(s.diag), len(s.ij), _p(s.ij),
                                               _p(s.blocks), size))

    def __del__(self):
        if getattr(self, "h", None):
            lib().mor_free(self.h)
            self.h = None

    def solve(self, system, b, x0=None, tol=1e-3):
        """(x, rounds) with rounds None where the reference returns None."""
        s = system
        b = np.ascontiguousarray(b, np.float64).reshape(-1, 3)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, np.float64).reshape(-1, 3).copy()
        r = lib().mor_solve(self.h, s.n, _p(s.diag), len(s.ij), _p(s.ij), _p(s.blocks), _p(b), _p(x), tol)
        return x, (None if r < 0 else int(r))

    def precondition(self, system, b):
        s = system
        b = np.ascontiguousarray(b, np.float64).reshape(-1, 3)
        z = np.zeros_like(b)
        lib().mor_precondition(self.h, s.n, _p(s.diag), len(s.ij), _p(s.ij), _p(s.blocks), _p(b), _p(z))
        return z

    def levels(self, cap=64):
        """[(parts, pairs, step, direct)] per coarse level, and the finest level's step."""
        parts, npairs = np.zeros(cap, np.int64), np.zeros(cap, np.int64)
        steps, direct, step0 = np.zeros(cap), np.zeros(cap, np.int32), ctypes.c_double()
        n = lib().mor_levels(self.h, cap, _p(parts), _p(npairs), _p(steps), _p(direct), ctypes.byref(step0))
        return [(int(parts[l]), int(npairs[l]), float(steps[l]), bool(direct[l])) for l in range(n)], step0.value

    def level_map(self, l, n_finer):
        of, r = np.zeros(n_finer, np.uint32), np.zeros((n_finer, 3))
        lib().mor_level_map(self.h, l, _p(of), _p(r))
        return of, r


def mech_lift(m, r, s):
    m = np.ascontiguousarray(m, np.float64)
    out = np.zeros((6, 6))
    lib().mor_lift(_p(m), _p(np.asarray(r, np.float64)), _p(np.asarray(s, np.float64)), _p(out))
    return out


def mech_inverse6(m):
    out = np.zeros((6, 6))
    lib().mor_inverse6(_p(np.ascontiguousarray(m, np.float64)), _p(out))
    return out


def mech_unit_k():
    """The unit cube's stiffness (mechanics/src/element.rs:49-76): (8, 8, 3, 3)."""
    k = np.zeros((8, 8, 3, 3))
    lib().mor_unit_k(_p(k))
    return k
