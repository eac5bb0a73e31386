"""The mechanics systems of the tests: parallax-world_amd/mech_systems.py (the engine's
assembly, mechanics/src/lib.rs:638-676) with the shapes the tests use."""
import mech_systems
import oracle as orc
from mech_systems import CORNERS, unit_k  # noqa: F401


def system(*args, **kw):
    """mech_systems.system, the system as the oracle's MechSystem (times, dense)."""
    s, b, nn, keys = mech_systems.system(*args, **kw)
    return orc.MechSystem(s.sites, s.at, s.diag, s.ij, s.blocks), b, nn, keys


def post(height, width=2, foot_held=True):
    """A post of width^2 x height cells standing on a held ground layer (the nodes at y = 0 held)."""
    import numpy as np
    sites = np.array([(x, y, z) for x in range(width) for y in range(height) for z in range(width)])
    return sites, (lambda k: k[:, 1] == 0) if foot_held else (lambda k: np.zeros(len(k), bool))


def blob(n, fill=0.5, seed=0):
    """Cells of an n^3 box kept with probability `fill`, plus a full bottom layer; nodes at y = 0 held."""
    import numpy as np
    rng = np.random.default_rng(seed)
    g = np.argwhere(np.ones((n, n, n), bool))
    keep = (rng.random(len(g)) < fill) | (g[:, 1] == 0)
    return g[keep], (lambda k: k[:, 1] == 0)
