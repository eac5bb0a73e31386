"""The benchmark configurations, each on the device through the C ABI.

* Config 1 (one structure, lognormal stiffness sigma 0.5, gravity + 64 point loads of
  magnitude U(1, 10)): FP64 and mixed solves to the configuration's 1e-8 against the
  oracle's 1e-13 solve at 44^3; at the full 256^3, by properties: both precisions reach a
  true residual of 1e-8 and agree with an FP64 device solve at 1e-12 to 1e-6.
* Config 2 (a batch of uneven structures, mixed): one st_solve_batch call over 16
  structures of side 16-48 and fill 15-40%, each structure against its own oracle solve.
* Config 4 (a sparse world, 15% solid, gravity + 1024 point loads): the mixed forward solve
  and both adjoints at 64^3 against the oracle's solve and the oracle's central differences
  at 8 cells; at the full 512^3, by properties: both solves reach 1e-8 and
  sum dC/dtheta = -C.

Bars: displacements within 1e-6 relative (max norm) of the tightly converged oracle solve;
gradients within 1e-4 of central differences (eps = 1e-4), relative to
max(|fd|, 1e-3 max |dJ|), as the other gradient tests.
"""
import numpy as np
import pytest

import oracle as orc
import statics_hip as sh
from harness import Pair, device_loads, generated, mats_from, point_loads

pytestmark = pytest.mark.gpu

WIDTHS = dict(pillar_min=2, pillar_max=6)  # pillars 2-6 cells wide, as the configurations state


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def classify_both(P):
    jo = P.L.joined_to_held()
    jg = P.from_dev(P.ctx.joined_to_held())  # also sets gone on the device
    assert (jo == jg).all()
    return (jo == 0).astype(np.uint8)


def softmax_J(L, x, beta, live):
    f = L.moved(x)
    sigma = ((f * L.n).sum(axis=1) / L.a)[live]
    m = sigma.max()
    return m + np.log(np.exp(beta * (sigma - m)).sum()) / beta


@pytest.mark.timeout(600)
def test_config1_point_loads_parity(ctx):
    """Config 1's workload at 44^3 (fill ~29% at pillar spacing 7): FP64 and mixed to the
    configuration's tolerance 1e-8, true residual, against the oracle at 1e-13."""
    P = generated(ctx, 44, 21, sigma=0.5, pillar_spacing=7.0, **WIDTHS)
    gone = classify_both(P)
    f = point_loads(P, gone, np.random.default_rng(21), 64)
    x_or, st_or = P.L.motions(f, gone=gone, tol=1e-13, max_iter=5000)
    assert not st_or["capped"]
    for prec in (sh.FP64, sh.MIXED, sh.MIXED32):
        x_dev, st = ctx.solve(P.to_dev(f), precision=prec, tol=1e-8, max_iter=2000)
        assert st.status == 0 and st.rel_residual <= 1e-8 and st.retries == 0
        err = rel(P.from_dev(x_dev), x_or)
        print(f"config 1 at 44^3, precision {prec}: {st.iters} iterations, residual {st.rel_residual:.2e}, "
              f"error {err:.2e}")
        assert err <= 1e-6


@pytest.mark.timeout(600)
def test_config2_batch_parity(ctx):
    """Config 2's batch rule (side 16 * 3^U, fill U(0.15, 0.40) through the pillar spacing,
    lognormal stiffness, gravity) for 16 structures in one mixed st_solve_batch call: every
    status OK and every structure within 1e-6 of its own oracle solve."""
    rng = np.random.default_rng(7)
    sets, sides, fills = [], [], []
    for i in range(16):
        side = int(16 * 3 ** rng.uniform(0.0, 1.0))
        phi = rng.uniform(0.15, 0.40)
        occ, theta = sh.generate(side, side, side, seed=700 + i, sigma=0.5,
                                 pillar_spacing=6.5 * np.sqrt(0.32 / phi), **WIDTHS)
        sets.append(sh.cells_from_occ(occ, theta))
        sides.append(side)
        fills.append(float((occ > 0).mean()))
    assert len(set(sides)) > 8 and min(fills) < 0.2 and max(fills) > 0.3
    xs, sts = ctx.solve_batch(sets, precision=sh.MIXED, tol=1e-8, max_iter=2000)
    assert all(s.status == 0 for s in sts), [s.status for s in sts]
    assert all(s.rel_residual <= 1e-8 for s in sts)
    worst = 0.0
    for (sites, held, e, g), xb in zip(sets, xs):
        P = Pair(ctx, sites, mats_from(e, g), held)
        gone = classify_both(P)
        f = np.zeros((P.L.n_nodes, 6))
        f[:, 2] = -P.L.m / orc.UNIT ** 3
        x_or, _ = P.L.motions(f, gone=gone, tol=1e-12, max_iter=5000)
        worst = max(worst, rel(P.from_dev(xb), x_or))
    print(f"config 2: 16 structures, sides {min(sides)}..{max(sides)}, fill {min(fills):.2f}..{max(fills):.2f}, "
          f"iterations {min(s.iters for s in sts)}..{max(s.iters for s in sts)}, worst error {worst:.2e}")
    assert worst <= 1e-6


@pytest.mark.timeout(900)
def test_config4_sparse_world_parity(ctx):
    """Config 4's workload at 64^3 (pillar spacing 12.8: ~15% solid, 1024 point loads): the
    mixed forward solve, then the compliance and soft-max (beta 20) adjoints with the
    hierarchy recomputed, against the oracle. Central differences of the oracle: for each of
    8 free cells and each sign, d solves K(theta +- eps) d = -(K(theta +- eps) - K) x, so
    x(theta +- eps) = x + d and C(+-) - C = f.d keep their digits."""
    P = generated(ctx, 64, 41, sigma=0.5, pillar_spacing=12.8, **WIDTHS)
    fill = len(P.sites) / 64 ** 3
    assert 0.12 < fill < 0.18
    gone = classify_both(P)
    f = point_loads(P, gone, np.random.default_rng(41), 1024)
    beta = 20.0
    x_or, _ = P.L.motions(f, gone=gone, tol=1e-13, max_iter=5000)
    x_dev, st = ctx.solve(P.to_dev(f), precision=sh.MIXED, tol=1e-8, max_iter=2000, release_hierarchy=True)
    assert st.status == 0 and st.rel_residual <= 1e-8
    err = rel(P.from_dev(x_dev), x_or)
    assert err <= 1e-6
    C, dC, sc = ctx.backward(sh.COMPLIANCE, precision=sh.MIXED, tol=1e-8)
    S, dS, ss = ctx.backward(sh.SOFTMAX_STRESS, beta=beta, precision=sh.MIXED, tol=1e-8, release_hierarchy=True)
    assert ss.status == 0 and ss.rel_residual <= 1e-8
    assert ss.setup_ms > 0  # the forward solve freed its hierarchy: the adjoint built it again
    live = ~(gone[P.L.ij[:, 0]].astype(bool) | gone[P.L.ij[:, 1]].astype(bool))
    assert abs(C - (f * x_or).sum()) <= 1e-7 * abs(C)
    assert abs(S - softmax_J(P.L, x_or, beta, live)) <= 1e-7 * abs(S)
    y0 = P.L.apply(x_or, gone=gone)
    rng = np.random.default_rng(42)
    eps = 1e-4
    worst = {"compliance": 0.0, "softmax": 0.0}
    for c in rng.choice(np.flatnonzero(P.held_cells == 0), size=8, replace=False):
        Jc, Js = [], []
        for sgn in (1.0, -1.0):
            Lq = orc.Lattice(0, P.sites, P.mats, P.held_cells)
            Lq.scale_cell(c, np.exp(sgn * eps))
            b = y0 - Lq.apply(x_or, gone=gone)
            d, _ = Lq.motions(b, gone=gone, tol=1e-9, max_iter=5000)
            Jc.append((f * d).sum())
            Js.append(softmax_J(Lq, x_or + d, beta, live))
        for name, dJ, pair in (("compliance", dC, Jc), ("softmax", dS, Js)):
            fd = (pair[0] - pair[1]) / (2 * eps)
            e = abs(dJ[c] - fd) / max(abs(fd), 1e-3 * np.abs(dJ).max())
            worst[name] = max(worst[name], e)
    print(f"config 4 at 64^3 (fill {fill:.3f}): forward {st.iters} iterations, error {err:.2e}; adjoint "
          f"{ss.iters} iterations (setup {ss.setup_ms:.0f} ms); gradient errors {worst}")
    assert worst["compliance"] <= 1e-4 and worst["softmax"] <= 1e-4


@pytest.mark.timeout(600)
def test_config1_full_size_properties(ctx):
    """Config 1 at its full 256^3: the mixed and FP64 solves to 1e-8 (true residual) agree with
    an FP64 device solve at 1e-12 to within 1e-6."""
    n = 256
    occ, theta = sh.generate(n, n, n, seed=1, sigma=0.5, pillar_spacing=7.0, **WIDTHS)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    ctx.upload(sites, held, e, g)
    joined = ctx.joined_to_held()
    f = device_loads(ctx, occ, joined, np.random.default_rng(1), 64)
    del occ, theta
    print(f"config 1 at 256^3: {ctx.nodes} nodes, {ctx.cells} free cells uploaded", flush=True)
    # the reference solution: FP64 as far as its true residual goes within a bounded count (the
    # FP64 true residual of this structure floors near 4e-10, 20 times below the tolerance)
    x_ref, s_ref = ctx.solve(f, precision=sh.FP64, tol=1e-12, max_iter=400, check=False)
    print(f"config 1 at 256^3: FP64 reference {s_ref.iters} iterations, residual {s_ref.rel_residual:.2e}",
          flush=True)
    assert s_ref.rel_residual <= 1e-9
    for prec in (sh.MIXED, sh.MIXED32, sh.FP64):
        x, st = ctx.solve(f, precision=prec, tol=1e-8, max_iter=2000)
        err = rel(x, x_ref)
        print(f"config 1 at 256^3, precision {prec}: {st.iters} iterations, residual {st.rel_residual:.2e}, "
              f"error against FP64 at 1e-12: {err:.2e}")
        assert st.status == 0 and st.rel_residual <= 1e-8 and err <= 1e-6


@pytest.mark.timeout(900)
def test_config4_full_size_properties(ctx):
    """Config 4 at its full 512^3 (occupied bricks only): the mixed forward solve and the mixed
    soft-max adjoint with the hierarchy recomputed both reach 1e-8, and the compliance
    gradient sums to -C (theta -> theta + t scales x by exp(-t))."""
    n = 512
    occ, theta = sh.generate(n, n, n, seed=4, sigma=0.5, pillar_spacing=12.8, **WIDTHS)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    del theta
    ctx.upload(sites, held, e, g)
    joined = ctx.joined_to_held()
    f = device_loads(ctx, occ, joined, np.random.default_rng(4), 1024)
    fill = len(sites) / n ** 3
    del occ, sites, held, e, g
    print(f"config 4 at 512^3: {ctx.nodes} nodes, {ctx.cells} free cells uploaded", flush=True)
    x, st = ctx.solve(f, precision=sh.MIXED, tol=1e-8, max_iter=2000, release_hierarchy=True)
    print(f"config 4 at 512^3: forward {st.iters} iterations, residual {st.rel_residual:.2e}", flush=True)
    assert st.status == 0 and st.rel_residual <= 1e-8
    C, dC, _ = ctx.backward(sh.COMPLIANCE, precision=sh.MIXED, tol=1e-8)
    assert abs(dC.sum() + C) <= 1e-7 * abs(C)
    S, dS, sa = ctx.backward(sh.SOFTMAX_STRESS, beta=20.0, precision=sh.MIXED, tol=1e-8, max_iter=2000,
                             release_hierarchy=True)
    assert sa.status == 0 and sa.rel_residual <= 1e-8 and sa.setup_ms > 0
    print(f"config 4 at 512^3 (fill {fill:.3f}, {st.dof_nodes * 6} DOF): forward {st.iters} iterations, "
          f"adjoint {sa.iters}; sum dC + C = {dC.sum() + C:.3e} (C {C:.6g}); device memory peak "
          f"{max(st.mem_peak, sa.mem_peak) / 1e9:.1f} GB")
