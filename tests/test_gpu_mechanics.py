"""mechanics::solve::Solver on the device (include/mechanics_hip.h) against the oracle's
restatement of mechanics/src/solve.rs: the same levels (parts and pairs per level, which is
solved directly), the same smoothing steps to rounding, the same rounds and solutions, on
systems built as mechanics/src/lib.rs:638-676 builds them (tests/mech_harness.py)."""
import numpy as np
import pytest

import mech_harness as mh
import oracle as orc
import statics_hip as sh

pytestmark = pytest.mark.gpu

H = 0.125


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


CASES = {
    "post 2x8x2": lambda: mh.post(8),
    "blob 8^3, 60%": lambda: mh.blob(8, 0.6, 1),
    "blob 16^3, 45%": lambda: mh.blob(16, 0.45, 2),
    "two posts apart": lambda: (np.concatenate([mh.post(6)[0], mh.post(6)[0] + [5, 0, 0]]), lambda k: k[:, 1] == 0),
}


def make(case, turn=0.3, seed=3):
    sites, held = CASES[case]()
    return mh.system(sites, held, h=H, turn=turn, seed=seed)


@pytest.mark.parametrize("case", list(CASES))
def test_levels_and_solve_match_oracle(case):
    s, b, _, _ = make(case)
    dev = sh.MechSolver(s, H)
    orc_s = orc.MechSolver(s, H)
    info = dev.info()
    lv, step0 = orc_s.levels()
    print(f"{case}: {s.n} unknowns, {len(s.ij)} pairs; levels {[(p, q, d) for p, q, _, d in info['levels']]}; "
          f"setup {info['setup_ms']:.1f} ms")
    assert [(p, q, d) for p, q, _, d in info["levels"]] == [(p, q, d) for p, q, _, d in lv]
    assert abs(info["step0"] - step0) <= 1e-9 * step0
    for (_, _, sd, _), (_, _, so, _) in zip(info["levels"], lv):
        assert abs(sd - so) <= 1e-9 * so
    # tightly, against the oracle and the dense solve
    x_or, r_or = orc_s.solve(s, b, tol=1e-10)
    x_dev, r_dev = dev.solve(s, b, tol=1e-10)
    xd = np.linalg.solve(s.dense(), b.reshape(-1)).reshape(-1, 3)
    print(f"  tol 1e-10: rounds {r_dev} (oracle {r_or}), error {rel(x_dev, x_or):.2e} against the oracle, "
          f"{rel(x_dev, xd):.2e} against the dense solve")
    assert r_dev is not None and abs(r_dev - r_or) <= 1
    assert rel(x_dev, x_or) < 1e-8 and rel(x_dev, xd) < 1e-8
    # at the engine's tolerance (mechanics/src/lib.rs:67), from a start
    x0 = np.random.default_rng(1).normal(size=b.shape) * 1e-3
    x_or, r_or = orc_s.solve(s, b, x0=x0, tol=1e-3)
    x_dev, r_dev = dev.solve(s, b, x0=x0, tol=1e-3)
    print(f"  tol 1e-3: rounds {r_dev} (oracle {r_or}), difference {rel(x_dev, x_or):.2e}")
    assert r_dev == r_or and rel(x_dev, x_or) < 1e-9


def test_levels_serve_a_system_close_to_theirs():
    """Levels made for one system solve another of the same unknowns (mechanics/src/lib.rs:741:
    the cells a little turned, a contact's pairs added), as the oracle's do."""
    s, b, _, keys = make("blob 8^3, 60%", turn=0.2, seed=4)
    s2, b2, _, keys2 = make("blob 8^3, 60%", turn=0.35, seed=5)
    assert (keys == keys2).all()
    # a contact between two far unknowns: a pair twice is the sum of its blocks
    extra = np.array([[0, s2.n - 1], [0, s2.n - 1]], np.uint32)
    kb = np.eye(3) * 1e3
    s2 = orc.MechSystem(s2.sites, s2.at, s2.diag + 0.0, np.concatenate([s2.ij, extra]),
                        np.concatenate([s2.blocks, -np.stack([kb, kb])]))
    s2.diag[0] += 2 * kb
    s2.diag[s2.n - 1] += 2 * kb
    dev, o = sh.MechSolver(s, H), orc.MechSolver(s, H)
    for tol in (1e-3, 1e-10):
        x_or, r_or = o.solve(s2, b2, tol=tol)
        x_dev, r_dev = dev.solve(s2, b2, tol=tol)
        print(f"  another system at {tol}: rounds {r_dev} (oracle {r_or}), difference {rel(x_dev, x_or):.2e}")
        assert r_dev is not None and abs(r_dev - r_or) <= (0 if tol > 1e-6 else 1) and rel(x_dev, x_or) < 1e-8
    # and back to its own
    x_dev, r_dev = dev.solve(s, b, tol=1e-10)
    x_or, r_or = o.solve(s, b, tol=1e-10)
    assert rel(x_dev, x_or) < 1e-8


def test_none_and_empty():
    """None where the reference returns None (a NaN measure), x left as it was; an empty system
    solves in no rounds; a system of other unknowns is refused."""
    s, b, _, _ = make("post 2x8x2")
    dev = sh.MechSolver(s, H)
    bad = b.copy()
    bad[3, 1] = np.nan
    x0 = np.ones_like(b)
    x, r = dev.solve(s, bad, x0=x0, tol=1e-3)
    _, r_or = orc.MechSolver(s, H).solve(s, bad, x0=x0, tol=1e-3)
    assert r is None and r_or is None and (x == x0).all()
    e = orc.MechSystem(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros((0, 3, 3)), np.zeros((0, 2)), np.zeros((0, 3, 3)))
    x, r = sh.MechSolver(e, H).solve(e, np.zeros((0, 3)))
    assert r == 0 and x.shape == (0, 3)
    s2, b2, _, _ = make("blob 8^3, 60%")
    with pytest.raises(sh.StaticsError) as ex:
        dev.solve(s2, b2)
    assert ex.value.code == sh.ST_BAD_INPUT
