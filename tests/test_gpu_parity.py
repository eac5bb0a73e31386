"""Device path against the oracle, through the C ABI.

Bars: classification (joined_to_held) bit-exact; the operator and bond forces to FP64
rounding; displacements within 1e-6 relative of the oracle's tightly converged solve
(the parity bar for solves; the tests ask for 1e-8 at tol 1e-10); gradients within 1e-4 of
central differences (eps = 1e-4) of the oracle's solve.
"""
import numpy as np
import pytest

import oracle as orc
import scenes
import statics_hip as sh
from harness import Pair, generated, mats_from

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def classify_both(P):
    jo = P.L.joined_to_held()
    jg = P.from_dev(P.ctx.joined_to_held())  # also sets gone on the device
    assert (jo == jg).all()
    return (jo == 0).astype(np.uint8)


def gravity_z(L):
    f = np.zeros((L.n_nodes, 6))
    f[:, 2] = -L.m / orc.UNIT ** 3
    return f


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_apply_parity_generated(ctx, seed):
    P = generated(ctx, 24, seed)
    gone = classify_both(P)
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(P.L.n_nodes, 6))
    y_or = P.L.apply(x, gone=gone)
    y_dev = P.from_dev(ctx.apply(P.to_dev(x)))
    assert rel(y_dev, y_or) < 1e-13


def test_apply_parity_every_schedule(ctx, monkeypatch):
    """Each schedule of the operator (node by node with and without LDS tiles, cell by cell;
    XCD brick mapping on and off, workgroup sizes) against the oracle, with dead beams and a
    gone set."""
    monkeypatch.setenv("STATICS_HIP_SCHEDULE", "999")  # clamped to the last schedule
    P = generated(ctx, 12, 4)
    count = ctx.apply_schedule(True)[0] + 1
    assert count >= 6
    for k in range(count):
        monkeypatch.setenv("STATICS_HIP_SCHEDULE", str(k))
        P = generated(ctx, 20, 4)
        rng = np.random.default_rng(k)
        beam_alive = rng.random(P.L.n_bonds // 4) > 0.1
        alive = np.repeat(beam_alive, 4).astype(np.uint8)
        jo = P.L.joined_to_held(alive)
        jg = P.from_dev(ctx.joined_to_held(P.beam_masks(alive.astype(bool))))
        assert (jo == jg).all()
        x = rng.normal(size=(P.L.n_nodes, 6))
        y = P.from_dev(ctx.apply(P.to_dev(x)))
        assert rel(y, P.L.apply(x, alive=alive, gone=(jo == 0).astype(np.uint8))) < 1e-13, k
        f = gravity_z(P.L)
        x_dev, st = ctx.solve(P.to_dev(f), precision=sh.MIXED, tol=1e-10)
        x_or, _ = P.L.motions(f, alive=alive, gone=(jo == 0).astype(np.uint8), tol=1e-13, max_iter=5000)
        assert st.status == 0 and rel(P.from_dev(x_dev), x_or) < 1e-8, k
        for fp64 in (True, False):
            idx, kernel, desc, disp = ctx.apply_schedule(fp64)
            cells = desc.startswith("cell-centric")
            assert idx == k and kernel.startswith("k_apply_cells" if cells else "k_apply"), (k, kernel)
            assert disp in ((1, 2) if cells else (1,)) and "threads a brick" in desc, (k, desc, disp)


def test_hierarchy_kept_and_dropped(ctx):
    """A second solve of the same operator keeps the hierarchy (same iterations and solution,
    setup down to bookkeeping); st_set_stiffness drops it, so the solve after it matches one
    after a fresh upload of the new stiffness; a change of precision mode rebuilds too."""
    occ, theta = sh.generate(32, 32, 32, seed=4, sigma=0.5)
    sites, held, e, g = sh.cells_from_occ(occ, theta)

    def fresh(e_, g_):
        ctx.upload(sites, held, e_, g_)
        ctx.joined_to_held()
        f = np.zeros((ctx.nodes, 6))
        f[:, 2] = -1.0
        f[ctx.node_held().astype(bool)] = 0
        return f

    f = fresh(e, g)
    x1, s1 = ctx.solve(f, precision=sh.MIXED, tol=1e-10)
    x2, s2 = ctx.solve(f, precision=sh.MIXED, tol=1e-10)
    assert s1.status == 0 and s2.status == 0 and s2.iters == s1.iters
    assert rel(x2, x1) < 1e-13 and s2.setup_ms < 0.5 * s1.setup_ms, (s1.setup_ms, s2.setup_ms)
    # stiffen the upper half a hundredfold: a kept hierarchy would no longer fit the operator
    scale = np.where(sites[:, 2] >= 16, 100.0, 1.0)
    ctx.set_stiffness(e * scale, g * scale)
    x3, s3 = ctx.solve(f, precision=sh.MIXED, tol=1e-10)
    f4 = fresh(e * scale, g * scale)
    assert (f4 == f).all()
    x4, s4 = ctx.solve(f, precision=sh.MIXED, tol=1e-10)
    assert s3.status == 0 and s3.iters == s4.iters and rel(x3, x4) < 1e-12, (s3.iters, s4.iters)
    x5, s5 = ctx.solve(f, precision=sh.FP64, tol=1e-10)
    assert s5.status == 0 and rel(x5, x4) < 1e-8 and s5.setup_ms > 2 * s2.setup_ms


def test_hierarchy_recomputed(ctx):
    """release_hierarchy: the solve frees its hierarchy, so the next solve and the adjoint build
    it again (setup well above the kept solve's bookkeeping) and give the same results as with
    the hierarchy kept (a rebuilt hierarchy agrees to rounding, so to the tolerance)."""
    occ, theta = sh.generate(32, 32, 32, seed=6, sigma=0.5)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    ctx.upload(sites, held, e, g)
    ctx.joined_to_held()
    f = np.zeros((ctx.nodes, 6))
    f[:, 2] = -1.0
    f[ctx.node_held().astype(bool)] = 0
    x1, s1 = ctx.solve(f, precision=sh.MIXED, tol=1e-10)
    J1, d1, a1 = ctx.backward(sh.SOFTMAX_STRESS, beta=20.0, precision=sh.MIXED, tol=1e-10)
    assert a1.setup_ms < 0.5 * s1.setup_ms  # kept: bookkeeping only
    x2, s2 = ctx.solve(f, precision=sh.MIXED, tol=1e-10, release_hierarchy=True)
    assert s2.setup_ms < 0.5 * s1.setup_ms  # this solve still found the kept one
    J2, d2, a2 = ctx.backward(sh.SOFTMAX_STRESS, beta=20.0, precision=sh.MIXED, tol=1e-10, release_hierarchy=True)
    assert a2.setup_ms > 2 * a1.setup_ms and a2.status == 0  # recomputed
    x3, s3 = ctx.solve(f, precision=sh.MIXED, tol=1e-10)
    assert s3.setup_ms > 2 * s2.setup_ms  # the adjoint released it again
    assert rel(x2, x1) < 1e-12 and rel(x3, x1) < 1e-8
    assert abs(J2 - J1) <= 1e-12 * abs(J1) and rel(d2, d1) < 1e-7
    assert s1.mem_peak > 0 and s1.retries == 0


def test_apply_parity_dead_beams_and_gone(ctx):
    P = generated(ctx, 16, 7)
    rng = np.random.default_rng(7)
    beam_alive = rng.random(P.L.n_bonds // 4) > 0.2
    alive = np.repeat(beam_alive, 4).astype(np.uint8)
    jo = P.L.joined_to_held(alive)
    jg = P.from_dev(ctx.joined_to_held(P.beam_masks(alive.astype(bool))))
    assert (jo == jg).all()
    gone = (jo == 0).astype(np.uint8)
    x = rng.normal(size=(P.L.n_nodes, 6))
    assert rel(P.from_dev(ctx.apply(P.to_dev(x))), P.L.apply(x, alive=alive, gone=gone)) < 1e-13
    # an arbitrary gone set: beams touching it are not live
    gone2 = (rng.random(P.L.n_nodes) < 0.05).astype(np.uint8)
    ctx.set_beams(None)
    ctx.set_gone(P.to_dev(gone2.astype(float)).astype(np.uint8))
    assert rel(P.from_dev(ctx.apply(P.to_dev(x))), P.L.apply(x, gone=gone2)) < 1e-13


@pytest.mark.parametrize("precision", [sh.FP64, sh.MIXED, sh.MIXED32])
@pytest.mark.parametrize("seed", [0, 1])
def test_solve_parity_generated(ctx, precision, seed):
    P = generated(ctx, 32, seed)
    gone = classify_both(P)
    f = gravity_z(P.L)
    x_or, st_or = P.L.motions(f, gone=gone, tol=1e-13, max_iter=5000)
    x_dev, st = ctx.solve(P.to_dev(f), precision=precision, tol=1e-10)
    assert st.status == 0 and st.rel_residual <= 1e-10
    assert rel(P.from_dev(x_dev), x_or) < 1e-8


@pytest.mark.parametrize("scene", [scenes.cantilever(orc.STONE, 0.5, 3.0), scenes.portico(1, 0.5)])
def test_solve_parity_reference_scenes(ctx, scene):
    sites, mats, held = scenes.solids(0, *scene)
    P = Pair(ctx, sites, mats, held)
    gone = classify_both(P)
    f = P.L.gravity()  # -y, as forces (statics/src/bodies.rs:408)
    x_or, _ = P.L.motions(f, gone=gone, tol=1e-13, max_iter=5000)
    # stone's stiff slab floors the FP64 true residual near 5e-10 for the oracle too (its recursive
    # residual reaches 1e-14): these scenes are held to the solve parity bar, tol 1e-8 and 1e-6
    for prec in (sh.FP64, sh.MIXED, sh.MIXED32):
        x_dev, st = ctx.solve(P.to_dev(f), precision=prec, tol=1e-8)
        assert st.status == 0
        assert rel(P.from_dev(x_dev), x_or) < 1e-6
    # bond forces and used, as forces/used_standing do
    fb_or = P.L.moved(x_or)
    fb_dev = P.bond_forces_dev(P.from_dev(x_dev))
    assert rel(fb_dev, fb_or) < 1e-5
    u_or = P.L.used(fb_or)
    u_dev = P.L.used(fb_dev)
    assert abs(u_dev.max() - u_or.max()) <= 1e-5 * u_or.max()


@pytest.mark.parametrize("levels", ["2", "8"])
@pytest.mark.parametrize("case", ["generated", "stone cantilever", "portico"])
def test_half_coarse_operators(ctx, monkeypatch, case, levels):
    """The mixed V-cycle's coarse operators stored in FP16 (block-scaled; by default on the two
    largest coarse levels, here also on every FP32 one) against the oracle: the solves meet
    the parity bar, on the stiff stone scenes too."""
    monkeypatch.setenv("STATICS_HIP_HALF", "1")
    monkeypatch.setenv("STATICS_HIP_HALF_LEVELS", levels)
    if case == "generated":
        P = generated(ctx, 32, 3)
        f = gravity_z(P.L)
    else:
        scene = scenes.cantilever(orc.STONE, 0.5, 3.0) if case == "stone cantilever" else scenes.portico(1, 0.5)
        P = Pair(ctx, *scenes.solids(0, *scene))
        f = P.L.gravity()
    gone = classify_both(P)
    x_or, _ = P.L.motions(f, gone=gone, tol=1e-13, max_iter=5000)
    x_dev, st = ctx.solve(P.to_dev(f), precision=sh.MIXED, tol=1e-8)
    print(f"half coarse operators on {levels} levels, {case}: {st.iters} iterations, restarts {st.retries}")
    assert st.status == 0 and rel(P.from_dev(x_dev), x_or) < 1e-6


def test_fused_restriction_matches(ctx, monkeypatch):
    """The level-1 restriction fused into the operator's epilogue against the separate kernel:
    the same sums in the same order per aggregate, compiled inside another kernel (rounding may
    differ in the last bits): the same iterations and the same solution to 1e-8."""
    monkeypatch.setenv("STATICS_HIP_SCHEDULE", "1")
    out = []
    for fuse in (True, False):
        if not fuse:
            monkeypatch.setenv("STATICS_HIP_NOFUSE", "1")
        P = generated(ctx, 24, 6)
        classify_both(P)
        for prec in (sh.FP64, sh.MIXED):
            x, st = ctx.solve(P.to_dev(gravity_z(P.L)), precision=prec, tol=1e-10)
            assert st.status == 0
            out.append((prec, st.iters, x))
    for (p1, i1, x1), (p2, i2, x2) in zip(out[:2], out[2:]):
        assert p1 == p2 and i1 == i2 and rel(x1, x2) < 1e-8, (p1, i1, i2, rel(x1, x2))


def _device_structure(ctx, n, seed):
    occ, theta = sh.generate(n, n, n, seed=seed, sigma=0.5)
    ctx.upload(*sh.cells_from_occ(occ, theta))
    ctx.joined_to_held()  # sets the gone nodes
    f = np.zeros((ctx.nodes, 6))
    f[:, 2] = -1.0
    return f


def test_galerkin_layouts_agree(ctx, monkeypatch):
    """The coarse Galerkin products with tasks by (i, j) pairs and grouped by column j (the same
    terms added in another order) against each other: the same iterations and the same solution
    to 1e-10, in both precisions. A 64^3 structure, whose hierarchy has five coarse levels, and
    each solve rebuilding it."""
    f = _device_structure(ctx, 64, 1)
    out = {}
    for pairs in ("1", "0"):
        monkeypatch.setenv("STATICS_HIP_GAL_PAIRS", pairs)
        for prec in (sh.FP64, sh.MIXED):
            x, st = ctx.solve(f, precision=prec, tol=1e-8, release_hierarchy=True)
            assert st.status == 0 and st.levels >= 4
            out[pairs, prec] = (st.iters, x)
    for prec in (sh.FP64, sh.MIXED):
        (i1, x1), (i0, x0) = out["1", prec], out["0", prec]
        assert i1 == i0 and rel(x0, x1) < 1e-10, (prec, i1, i0, rel(x0, x1))


def test_coarse_layouts_agree(ctx, monkeypatch):
    """Every layout of the coarse residual kernels forced on every level (STATICS_HIP_WIDE=k:
    on the FP16 levels the record kernel's lanes per row and blocks in flight, elsewhere the
    lane layouts) against the tuned choice: mixed solves to 1e-10 on a 48^3 structure reach
    the same solution to 1e-8, within two iterations."""
    f = _device_structure(ctx, 48, 2)
    x0, st0 = ctx.solve(f, precision=sh.MIXED, tol=1e-10, release_hierarchy=True)
    assert st0.status == 0
    for k in range(8):
        monkeypatch.setenv("STATICS_HIP_WIDE", str(k))
        x, st = ctx.solve(f, precision=sh.MIXED, tol=1e-10, release_hierarchy=True)
        assert st.status == 0 and abs(st.iters - st0.iters) <= 2 and rel(x, x0) < 1e-8, (k, st.iters, rel(x, x0))


def test_transfer_variants_agree(ctx, monkeypatch):
    """The transfers below large FP32 levels through FP32 prolongator records against the block
    arrays (STATICS_HIP_PREC=0), and the switch to FP64 folded into the restriction and
    prolongation against separate conversions (STATICS_HIP_FOLD=0): the same sums in another
    order or the same values, so mixed solves to 1e-10 on a 128^3 structure (level 2 above
    8,192 aggregates) reach the same solution to 1e-8 within one iteration. The records in
    FP16 (the default) change the preconditioner slightly: the same solution to 1e-7 within
    three iterations."""
    f = _device_structure(ctx, 128, 1)
    monkeypatch.setenv("STATICS_HIP_PHALF", "0")
    x0, st0 = ctx.solve(f, precision=sh.MIXED, tol=1e-10, release_hierarchy=True)
    assert st0.status == 0
    for var in ("STATICS_HIP_PREC", "STATICS_HIP_FOLD"):
        monkeypatch.setenv(var, "0")
        x, st = ctx.solve(f, precision=sh.MIXED, tol=1e-10, release_hierarchy=True)
        monkeypatch.delenv(var)
        assert st.status == 0 and abs(st.iters - st0.iters) <= 1 and rel(x, x0) < 1e-8, (var, st.iters, rel(x, x0))
    monkeypatch.delenv("STATICS_HIP_PHALF")
    x, st = ctx.solve(f, precision=sh.MIXED, tol=1e-10, release_hierarchy=True)
    assert st.status == 0 and abs(st.iters - st0.iters) <= 3 and rel(x, x0) < 1e-7, (st.iters, st0.iters, rel(x, x0))


def test_reference_stopping_rule(ctx):
    """refresh < 0: the recursive residual decides (bodies.rs:350-352); the solution still meets
    the bar and rel_residual reports the true residual."""
    P = generated(ctx, 24, 5)
    gone = classify_both(P)
    f = gravity_z(P.L)
    x_or, _ = P.L.motions(f, gone=gone, tol=1e-13, max_iter=5000)
    for prec in (sh.FP64, sh.MIXED):
        x_dev, st = ctx.solve(P.to_dev(f), precision=prec, tol=1e-8, refresh=-1)
        assert st.status == 0 and st.refreshes == 0 and st.rel_residual < 1e-7
        assert rel(P.from_dev(x_dev), x_or) < 1e-6
    # MIXED32 replaces its FP32 residual by the true one as it goes, under either rule
    x_dev, st = ctx.solve(P.to_dev(f), precision=sh.MIXED32, tol=1e-8, refresh=-1)
    assert st.status == 0 and st.refreshes >= 1 and st.rel_residual < 1e-7
    assert rel(P.from_dev(x_dev), x_or) < 1e-6


def test_bond_forces_parity(ctx):
    P = generated(ctx, 16, 3)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(P.L.n_nodes, 6))
    x[P.L.held.astype(bool)] = 0
    assert rel(P.bond_forces_dev(x), P.L.moved(x)) < 1e-12


def test_bond_used_parity(ctx):
    """st_bond_used against Bond::used of the oracle's bond forces, per-cell strengths."""
    P = generated(ctx, 16, 3)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(P.L.n_nodes, 6)) * 1e-2
    x[P.L.held.astype(bool)] = 0
    strength = rng.uniform(0.5, 2.0, size=(len(P.sites), 3))
    u_dev = ctx.bond_used(strength, P.to_dev(x))[P.beam_cell, P.beam_q].reshape(-1)
    u_or = P.L.used(P.L.moved(x), np.repeat(strength[P.beam_cell], 4, axis=0))
    assert u_or.max() > 1.0 and rel(u_dev, u_or) < 1e-12


def settle_both(P, tol=1e-6, **kw):
    """The oracle's settle and st_settle on the same cells (gravity along -y, bodies.rs:408).
    The oracle solves to 1e-10 (its recursive residual); the device to the reference's own
    settle tolerance, 1e-6 (bodies.rs:432), in the FP64 true residual: after the first breaks
    some rounds' systems floor the true residual near 3e-6 (the portico's second round)."""
    o = P.L.settle(tol=1e-10)
    gone, beams, rounds, st = P.ctx.settle(P.mats[:, 3:], P.to_dev(P.L.gravity()), tol=tol, **kw)
    print(f"settle: {rounds} rounds, {st.iters} iterations, restarts {st.retries}, flagged {st.first_status}, "
          f"residual {st.rel_residual:.2e}")
    assert st.status == 0
    gone = P.from_dev(gone)
    free = ~P.L.held.astype(bool)
    fallen = P.L.m[gone.astype(bool)].sum() / P.L.m[free].sum()
    return o, gone, beams, rounds, fallen


@pytest.mark.parametrize("name,scene,kw", [
    ("stone beam 4 m", scenes.cantilever(orc.STONE, 0.5, 4.0), dict(precision=sh.MIXED, warm_start=True)),
    ("stone beam 8 m", scenes.cantilever(orc.STONE, 0.5, 8.0), dict(precision=sh.FP64)),
    ("portico, two columns cut", scenes.portico(2, 0.5), dict(precision=sh.MIXED, warm_start=True)),
    ("portico at 1e-8, the reference's stopping rule", scenes.portico(2, 0.5),
     dict(precision=sh.MIXED, warm_start=True, tol=1e-8, refresh=-1)),
    ("tree cut 30 %", scenes.tree(8.0, -1.0, 0.3), dict(precision=sh.MIXED, warm_start=True)),
    ("tree cut 70 %", scenes.tree(8.0, -1.0, 0.7), dict(precision=sh.MIXED)),
])
def test_settle_parity_scenes(ctx, name, scene, kw):
    """bodies::settle on the reference's scenes (statics/examples/statics.rs): the same rounds,
    the same gone nodes (bit-exact) and the same fallen share as the oracle."""
    sites, mats, held = scenes.solids(0, *scene)
    P = Pair(ctx, sites, mats, held)
    o, gone, beams, rounds, fallen = settle_both(P, **kw)
    assert rounds == o["rounds"], (name, rounds, o["rounds"])
    assert (gone == o["gone"]).all(), (name, int(gone.sum()), int(o["gone"].sum()))
    assert abs(fallen - o["fallen"]) <= 1e-12, (name, fallen, o["fallen"])
    if rounds == 1:
        assert (beams[held == 0] == 0x7FFF).all()


def test_settle_parity_generated(ctx):
    """A generated structure whose strengths its most stressed bonds overuse: breaking cascades
    over several rounds, identically on both sides; the final masks kill whole beams only."""
    occ, theta = sh.generate(20, 20, 20, seed=5, sigma=0.5)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    P0 = Pair(ctx, sites, mats_from(e, g), held)
    u = P0.L.used_standing(tol=1e-10)
    s = float(np.quantile(u[u > 0], 0.995))
    P = Pair(ctx, sites, mats_from(e, g, strength=(s, s, s)), held)
    o, gone, beams, rounds, fallen = settle_both(P, precision=sh.MIXED, warm_start=True)
    assert o["rounds"] >= 2 and rounds == o["rounds"], (rounds, o["rounds"])
    assert (gone == o["gone"]).all() and abs(fallen - o["fallen"]) <= 1e-12
    # the context keeps the final masks: classifying with them gives the same gone set
    jg = P.from_dev(ctx.joined_to_held())
    assert ((jg == 0) == gone.astype(bool)).all()
    assert (beams[held == 0] != 0x7FFF).any()
    # every beam alive again: a second settle repeats the first (the loads set while most
    # nodes were gone still reach every node)
    ctx.set_beams(None)
    o2, gone2, beams2, rounds2, fallen2 = settle_both(P, precision=sh.MIXED, warm_start=True)
    assert rounds2 == rounds and (gone2 == gone).all() and (beams2 == beams).all()


def softmax_J(L, x, beta, live):
    f = L.moved(x)
    sigma = ((f * L.n).sum(axis=1) / L.a)[live]
    m = sigma.max()
    return m + np.log(np.exp(beta * (sigma - m)).sum()) / beta


@pytest.mark.parametrize("objective", [sh.COMPLIANCE, sh.SOFTMAX_STRESS])
def test_gradient_central_differences(ctx, objective):
    """dJ/dtheta against central differences of the oracle's solve, eps = 1e-4, solves to 1e-13."""
    P = generated(ctx, 12, 11, slab_min=4, slab_max=6)
    gone = classify_both(P)
    f = gravity_z(P.L)
    beta = 20.0
    x_dev, st = ctx.solve(P.to_dev(f), precision=sh.FP64, tol=1e-11)
    J, dJ, _ = ctx.backward(objective, beta=beta, precision=sh.FP64, tol=1e-11)
    rng = np.random.default_rng(11)
    free_cells = np.flatnonzero(P.held_cells == 0)
    eps = 1e-4

    live = ~(gone[P.L.ij[:, 0]].astype(bool) | gone[P.L.ij[:, 1]].astype(bool))

    def J_of(L):
        x, _ = L.motions(f, gone=gone, tol=1e-13, max_iter=5000)
        return (f * x).sum() if objective == sh.COMPLIANCE else softmax_J(L, x, beta, live)

    assert abs(J - J_of(P.L)) <= 1e-9 * abs(J)
    for c in rng.choice(free_cells, size=6, replace=False):
        Lp = orc.Lattice(0, P.sites, P.mats, P.held_cells)
        Lp.scale_cell(c, np.exp(eps))
        Lm = orc.Lattice(0, P.sites, P.mats, P.held_cells)
        Lm.scale_cell(c, np.exp(-eps))
        fd = (J_of(Lp) - J_of(Lm)) / (2 * eps)
        assert abs(dJ[c] - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(dJ).max()), (c, dJ[c], fd)


def test_gradient_device_differences(ctx):
    """dJ/dtheta against central differences of the device's own FP64 solves (tol 1e-11, near the
    FP64 floor of these systems) on a larger structure than the oracle's check: stiffness changed
    in place by st_set_stiffness."""
    occ, theta = sh.generate(40, 40, 40, seed=9, sigma=0.5)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    ctx.upload(sites, held, e, g)
    ctx.joined_to_held()
    nh = ctx.node_held()
    f = np.zeros((ctx.nodes, 6))
    f[:, 2] = -1.0
    f[nh.astype(bool)] = 0
    _, st = ctx.solve(f, precision=sh.FP64, tol=1e-11, max_iter=20000, check=False)
    assert st.rel_residual <= 1e-10
    J, dC, _ = ctx.backward(sh.COMPLIANCE, precision=sh.FP64, tol=1e-11, max_iter=20000)
    S, dS, _ = ctx.backward(sh.SOFTMAX_STRESS, beta=20.0, precision=sh.FP64, tol=1e-11, max_iter=20000)
    assert abs(ctx.objective(sh.COMPLIANCE) - J) <= 1e-12 * abs(J)
    assert abs(ctx.objective(sh.SOFTMAX_STRESS, 20.0) - S) <= 1e-12 * abs(S)
    rng = np.random.default_rng(9)
    eps = 1e-4
    for c in rng.choice(np.flatnonzero(held == 0), size=6, replace=False):
        Jp = []
        for sgn in (1.0, -1.0):
            ctx.set_stiffness(e * np.where(np.arange(len(e)) == c, np.exp(sgn * eps), 1.0),
                              g * np.where(np.arange(len(g)) == c, np.exp(sgn * eps), 1.0))
            _, st = ctx.solve(f, precision=sh.FP64, tol=1e-11, max_iter=20000, check=False)
            assert st.rel_residual <= 1e-10
            Jp.append((ctx.objective(sh.COMPLIANCE), ctx.objective(sh.SOFTMAX_STRESS, 20.0)))
        for k, d in ((0, dC), (1, dS)):
            fd = (Jp[0][k] - Jp[1][k]) / (2 * eps)
            assert abs(d[c] - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(d).max()), (k, c, d[c], fd)
    ctx.set_stiffness(e, g)


def test_gradient_scaling_identities(ctx):
    """theta -> theta + t scales x by exp(-t): sum dC/dtheta = -C and sum dJ_softmax/dtheta = 0."""
    P = generated(ctx, 48, 5)
    ctx.joined_to_held()
    f = gravity_z(P.L)
    ctx.solve(P.to_dev(f), precision=sh.MIXED, tol=1e-10)
    C, dC, _ = ctx.backward(sh.COMPLIANCE)
    assert abs(dC.sum() + C) <= 1e-8 * abs(C)
    J, dJ, _ = ctx.backward(sh.SOFTMAX_STRESS, beta=20.0, precision=sh.MIXED, tol=1e-10)
    assert abs(dJ.sum()) <= 1e-6 * np.abs(dJ).sum()


def test_batch_matches_single(ctx):
    sets = []
    for n, seed in ((16, 1), (24, 2), (20, 3)):
        occ, theta = sh.generate(n, n, n, seed=seed, sigma=0.5)
        sets.append(sh.cells_from_occ(occ, theta))
    # the batch classifies each structure and loads it with gravity (weight 1 per free cell, -z)
    xs_b, st_b = ctx.solve_batch([(s, h, e, g) for s, h, e, g in sets], precision=sh.FP64, tol=1e-10, max_iter=20000)
    for (s, h, e, g), xb, sb in zip(sets, xs_b, st_b):
        ctx.upload(s, h, e, g)
        ctx.joined_to_held()
        faces = ctx.node_faces()
        held = ctx.node_held()
        cnt = np.zeros(len(faces))
        occ_cells = {tuple(c) for c, hh in zip(s.tolist(), h) if not hh}
        for v, (o, x, y, z) in enumerate(faces):
            cnt[v] = ((x, y, z) in occ_cells) + ((x - (o == 0), y - (o == 1), z - (o == 2)) in occ_cells)
        f = np.zeros((len(faces), 6))
        f[:, 2] = -cnt / 6
        f[held.astype(bool)] = 0
        x1, s1 = ctx.solve(f, precision=sh.FP64, tol=1e-10, max_iter=20000, check=False)
        assert s1.status == sb.status
        if s1.status == 0:
            assert rel(xb, x1) < 1e-8


def test_batch_graph_replay(ctx, monkeypatch):
    """A batch's iterations replayed as captured graphs (eight iterations a launch) against the
    same batch launched kernel by kernel: the same kernels in the same order. Each batch call
    uploads and builds its hierarchy anew, and the Galerkin products' LDS atomics make a
    rebuilt hierarchy equal to rounding only, so the solutions agree to the solve's accuracy
    (1e-7 relative, the iteration counts within one); twice in a row, eager solves between."""
    sets = []
    for n, seed in ((16, 21), (28, 22), (20, 23), (36, 24)):
        occ, theta = sh.generate(n, n, n, seed=seed, sigma=0.5)
        sets.append(sh.cells_from_occ(occ, theta))
    monkeypatch.setenv("STATICS_HIP_GRAPHS", "0")
    xe, se = ctx.solve_batch(sets, precision=sh.MIXED, tol=1e-8, max_iter=2000)
    for rep in range(2):
        monkeypatch.setenv("STATICS_HIP_GRAPHS", "1")
        xg, sg = ctx.solve_batch(sets, precision=sh.MIXED, tol=1e-8, max_iter=2000)
        assert all(abs(a.iters - b.iters) <= 1 for a, b in zip(sg, se)) and all(s.status == 0 for s in sg)
        for a, b in zip(xg, xe):
            assert rel(a, b) < 1e-7
        monkeypatch.setenv("STATICS_HIP_GRAPHS", "0")
        xe2, _ = ctx.solve_batch(sets, precision=sh.MIXED, tol=1e-8, max_iter=2000)
        for a, b in zip(xe2, xe):
            assert rel(a, b) < 1e-7


def test_edge_cases(ctx):
    # one free cell on held ground
    sites = np.array([[0, 0, 0], [0, 0, 1]])
    held = np.array([1, 0], np.uint8)
    P = Pair(ctx, sites, mats_from(np.ones(2), 0.4 * np.ones(2)), held)
    f = gravity_z(P.L)
    x_or, _ = P.L.motions(f, tol=1e-13)
    x_dev, st = ctx.solve(P.to_dev(f), tol=1e-12)
    assert st.status == 0 and rel(P.from_dev(x_dev), x_or) < 1e-10
    # zero loads: no iteration, x = 0 (statics/src/bodies.rs:350-353)
    x0, st0 = ctx.solve(np.zeros((ctx.nodes, 6)), tol=1e-8)
    assert st0.status == 0 and st0.iters == 0 and not x0.any()
    # an iteration cap of 1 is reported, not hidden
    P = generated(ctx, 16, 0)
    ctx.joined_to_held()
    f = gravity_z(P.L)
    _, st1 = ctx.solve(P.to_dev(f), tol=1e-12, max_iter=1, check=False)
    assert st1.status == sh.ST_NO_CONVERGENCE
    # NaN loads are reported
    fn = P.to_dev(f)
    fn[np.flatnonzero(np.abs(fn).sum(axis=1) > 0)[0], 2] = np.nan
    _, st2 = ctx.solve(fn, tol=1e-8, check=False)
    assert st2.status == sh.ST_NAN
    # a floating piece (gone not set) is not SPD: the solve does not claim convergence
    ctx.set_gone(None)
    _, st3 = ctx.solve(P.to_dev(f), tol=1e-8, max_iter=300, check=False)
    assert st3.status != 0
    # bad input is refused
    with pytest.raises(sh.StaticsError):
        ctx.upload(np.array([[0, 0, 0], [0, 0, 0]]), np.zeros(2, np.uint8), np.ones(2), np.ones(2))
    with pytest.raises(sh.StaticsError):
        ctx.upload(np.array([[0, 0, 0]]), np.zeros(1, np.uint8), -np.ones(1), np.ones(1))


def grid_decode(min_site, n, bricks, solid, nsites):
    """Grid::solid over the box in solids' order (lattice/src/lib.rs:293-335; statics/src/lattice.rs:87-102)."""
    nb = [(v + 7) // 8 for v in n]
    sites, index = [], []
    for x in range(n[0]):
        for y in range(n[1]):
            for z in range(n[2]):
                at = bricks[((z >> 3) * nb[1] + (y >> 3)) * nb[0] + (x >> 3)]
                if at < 0:
                    continue
                i = at * 512 + ((z & 7) * 8 + (y & 7)) * 8 + (x & 7)
                if i < nsites and (solid[i >> 5] >> (i & 31)) & 1:
                    sites.append((x + min_site[0], y + min_site[1], z + min_site[2]))
                    index.append(i)
    return np.array(sites, np.int64).reshape(-1, 3), np.array(index, np.int64)


@pytest.mark.parametrize("level,radius", [(0, 0.2), (0, 0.3), (1, 0.5)])
def test_grid_cells_parity(ctx, level, radius):
    """st_grid_cells against Grid's own indexing and the oracle's hold_bulk (bit-exact): a box
    that is not a whole number of bricks, empty and shared bricks, an offset origin."""
    rng = np.random.default_rng(level * 10 + int(radius * 10))
    n = (37, 29, 45)
    min_site = (-5, 3, 10)
    nb = [(v + 7) // 8 for v in n]
    nstarts = 30
    bricks = rng.integers(0, nstarts, size=nb[0] * nb[1] * nb[2]).astype(np.int32)
    bricks[rng.random(len(bricks)) < 0.15] = -1
    nsites = nstarts * 512
    bits = (rng.random(nsites) < 0.9).astype(np.uint64)
    solid = np.zeros(nsites // 32, np.uint32)
    for k in range(32):
        solid |= (bits[k::32] << np.uint64(k)).astype(np.uint32)
    sites, held, index = ctx.grid_cells(min_site, n, bricks, solid, nsites, level=level, hold_radius=radius)
    want_sites, want_index = grid_decode(min_site, n, bricks, solid, nsites)
    assert (sites == want_sites).all() and (index == want_index).all()
    want_held = orc.hold_bulk(level, want_sites.astype(np.int64) << level, radius)
    assert (held == want_held).all() and 0 < held.sum() < len(held)


def test_voxelgrid_cells_parity(ctx):
    """The same for a worldgen VoxelGrid: a byte per site (how much of it stops a walking
    body), a cell from the threshold up (worldgen/src/grid.rs:530-537, 617-622)."""
    rng = np.random.default_rng(77)
    n = (30, 17, 41)
    min_site = (4, -9, 0)
    nb = [(v + 7) // 8 for v in n]
    nstarts = 24
    bricks = rng.integers(0, nstarts, size=nb[0] * nb[1] * nb[2]).astype(np.int32)
    bricks[rng.random(len(bricks)) < 0.1] = -1
    nsites = nstarts * 512
    sb = rng.integers(0, 256, size=nsites).astype(np.uint8)
    sb[rng.random(nsites) < 0.6] = 255
    thr = 200
    sites, held, index = ctx.grid_cells(min_site, n, bricks, sb, nsites, level=0, hold_radius=0.2, solid_min=thr)
    bits = (sb >= thr).astype(np.uint64)
    solid = np.zeros(nsites // 32, np.uint32)
    for k in range(32):
        solid |= (bits[k::32] << np.uint64(k)).astype(np.uint32)
    want_sites, want_index = grid_decode(min_site, n, bricks, solid, nsites)
    assert (sites == want_sites).all() and (index == want_index).all()
    assert (held == orc.hold_bulk(0, want_sites.astype(np.int64), 0.2)).all() and 0 < held.sum() < len(held)


def test_grid_cells_block():
    """A 3^3 block in one brick: with hold radius one cell only the centre is held; the cells
    upload and solve."""
    ctx = sh.Context()
    solid = np.zeros(512 // 32, np.uint32)
    for x in range(2, 5):
        for y in range(1, 4):
            for z in range(3, 6):
                i = (z * 8 + y) * 8 + x
                solid[i >> 5] |= np.uint32(1 << (i & 31))
    sites, held, index = ctx.grid_cells((0, 0, 0), (8, 8, 8), np.array([0], np.int32), solid, 512, 0, 0.125)
    assert len(sites) == 27 and held.sum() == 1 and tuple(sites[held == 1][0]) == (3, 2, 4)
    assert (sites[:, 0][:-1] <= sites[:, 0][1:]).all()  # x outermost
    e = np.ones(len(sites))
    ctx.upload(sites, held, e, 0.4 * e)
    ctx.joined_to_held()
    f = np.zeros((ctx.nodes, 6))
    f[:, 2] = -1.0
    f[ctx.node_held().astype(bool)] = 0
    x, st = ctx.solve(f, tol=1e-10)
    assert st.status == 0 and np.isfinite(x).all()


def test_settle_parity_collapse(ctx):
    """A generated structure under gravity along -y with every strength at the 0.999 quantile
    of its standing used shares: breaking cascades until it falls, round for round as in the
    oracle. The rounds where it comes apart are nearly singular systems (where the fast cycle
    settings may lose definiteness and the solve is repeated with the robust ones)."""
    occ, theta = sh.generate(32, 32, 32, seed=1, sigma=0.5, pillar_spacing=8.0, pillar_min=3, pillar_max=7)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    P0 = Pair(ctx, sites, mats_from(e, g), held)
    u = P0.L.used_standing(tol=1e-10)
    s = float(np.quantile(u[u > 0], 0.999))
    P = Pair(ctx, sites, mats_from(e, g, strength=(s, s, s)), held)
    o, gone, beams, rounds, fallen = settle_both(P, precision=sh.MIXED, warm_start=True, refresh=-1)
    assert o["rounds"] >= 3 and rounds == o["rounds"], (rounds, o["rounds"])
    assert (gone == o["gone"]).all() and abs(fallen - o["fallen"]) <= 1e-12


# ---------------------------------------------------------------- gather.rs

GATHER_FIELDS = ["x", "m", "node_mat", "held", "around", "ij", "p", "n", "a", "d", "bond_mat"]


def gather_same(dev, want):
    """The device's gathered lattice against the oracle's: the same nodes and bonds in the same
    order, every value equal (the measure runs the reference's operations in its order)."""
    assert dev.level == want.level
    assert len(dev.m) == want.n_nodes and len(dev.a) == want.n_bonds, (len(dev.m), want.n_nodes, len(dev.a),
                                                                       want.n_bonds)
    for f in GATHER_FIELDS:
        a, b = getattr(dev, f), getattr(want, f)
        b = b.reshape(a.shape)
        if not np.array_equal(a, b):
            diff = np.abs(a.astype(float) - b.astype(float))
            k = int(np.argmax(diff))
            raise AssertionError(f"{f}: {int((a != b).sum())} of {a.size} differ, worst at {k}: "
                                 f"{a.ravel()[k]!r} against {b.ravel()[k]!r}")


def _two_material_blob(n, seed):
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


@pytest.mark.parametrize("case", ["bar", "blob", "blob_measured", "truss", "cantilever"])
def test_gather_parity(ctx, case):
    """gather::gathered on the device equals the oracle's, level after level."""
    if case == "bar":
        fill = lambda x: (orc.STEEL, False) if (abs(x[0]) < 4.0 and abs(x[1]) < 0.2 and abs(x[2]) < 0.2) else None
        lat = orc.Lattice(0, *scenes.solids(0, [-4.5, -1.0, -1.0], [4.5, 1.0, 1.0], fill))
        levels, full = 3, True
    elif case.startswith("blob"):
        lat, levels, full = _two_material_blob(12, 3), 2, case == "blob"
    elif case == "truss":
        lat = orc.Lattice(0, np.array([[0, 0, 0], [1, 1, 1]]), [orc.STONE] * 2, np.zeros(2, np.uint8))
        levels, full = 1, True
    else:
        lat = orc.Lattice(0, *scenes.solids(0, *scenes.cantilever(orc.WOOD, 0.5, 3.0)))
        levels, full = 2, True
    print(f"gather {case}: {lat.n_nodes} nodes, {lat.n_bonds} bonds", flush=True)
    for _ in range(levels):
        want = lat.gathered(against_full=full)
        dev = ctx.gather(lat, against_full=full)
        gather_same(dev, want)
        lat = want


def test_gather_generated_structure(ctx):
    """A 32^3 structure of the bench's generator (pillars, slabs, holes; a held ground layer),
    gathered twice: equal to the oracle's, and its weight kept (gather.rs:400-412)."""
    occ, _ = sh.generate(32, 32, 32, seed=4, sigma=0.5, pillar_spacing=8.0, pillar_min=3, pillar_max=7)
    sites, held, _, _ = sh.cells_from_occ(occ)
    lat = orc.Lattice(0, sites.astype(np.int64), orc.MASONRY, held)
    whole = lat.m.sum()
    print(f"gather generated: {lat.n_nodes} nodes, {lat.n_bonds} bonds", flush=True)
    for _ in range(2):
        want = lat.gathered()
        dev = ctx.gather(lat)
        gather_same(dev, want)
        assert abs(dev.m.sum() / whole - 1.0) < 1e-12
        lat = want


def test_gather_empty_and_refused(ctx):
    """An empty lattice gathers to an empty one a level up; a bond naming a node that is not
    there, or a cell past the measure's limits, is refused with a message."""
    empty = sh.GatheredLattice(3, 0, 0)
    g = ctx.gather(empty)
    assert g.level == 4 and len(g.m) == 0 and len(g.a) == 0
    lat = orc.Lattice(0, np.array([[0, 0, 0]]), [orc.STONE], np.zeros(1, np.uint8))
    bad = sh.GatheredLattice(0, lat.n_nodes, lat.n_bonds)
    for f in GATHER_FIELDS:
        setattr(bad, f, getattr(lat, f).copy().reshape(getattr(bad, f).shape))
    bad.ij[5, 1] = lat.n_nodes
    with pytest.raises(RuntimeError, match="out of range"):
        ctx.gather(bad)
    # a bond whose nodes share no cell
    bad.ij[5, 1] = lat.ij[5, 1]
    bad.around[0] = [[100, 100, 100], [101, 100, 100]]
    with pytest.raises(RuntimeError, match="share no cell"):
        ctx.gather(bad)
    # more bonds in one coarse cell than any built or gathered lattice has (9 x 60 > 480)
    many = sh.GatheredLattice(0, lat.n_nodes, 9 * lat.n_bonds)
    for f in ["x", "m", "node_mat", "held", "around"]:
        setattr(many, f, getattr(lat, f).copy().reshape(getattr(many, f).shape))
    for f in ["ij", "p", "n", "a", "d", "bond_mat"]:
        v = getattr(lat, f).reshape(getattr(bad, f).shape)
        setattr(many, f, np.concatenate([v] * 9))
    with pytest.raises(RuntimeError, match="480 bonds"):
        ctx.gather(many)
    # and the context is still good for the next call
    gather_same(ctx.gather(lat), lat.gathered())


def test_gather_parity_two_launches(ctx):
    """More coarse cells than one launch measures (16,384): a 72^3 structure of the generator,
    gathered once, equal to the oracle's throughout."""
    occ, _ = sh.generate(72, 72, 72, seed=2, sigma=0.5, pillar_spacing=8.0, pillar_min=3, pillar_max=7)
    sites, held, _, _ = sh.cells_from_occ(occ)
    lat = orc.Lattice(0, sites.astype(np.int64), orc.STONE, held)
    # the coarse cells holding bonds: each bond's cell is the site its two nodes share
    i, j = lat.ij[::4, 0], lat.ij[::4, 1]
    ai, aj = lat.around[i], lat.around[j]
    first = (ai[:, 0] == aj[:, 0]).all(1) | (ai[:, 0] == aj[:, 1]).all(1)
    shared = np.where(first[:, None], ai[:, 0], ai[:, 1])
    cells = np.unique((shared >> 1) << 1, axis=0)
    assert len(cells) > 16384, len(cells)
    print(f"gather 72^3: {lat.n_nodes} nodes, {lat.n_bonds} bonds, {len(cells)} coarse cells with bonds", flush=True)
    gather_same(ctx.gather(lat), lat.gathered())
