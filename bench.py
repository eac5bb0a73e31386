#!/usr/bin/env python3
"""Structure solves/sec to 1e-8 relative residual at 256^3 on one MI355X.

Workload: one structure on a 256^3 lattice, ~29% solid: a held ground layer,
square pillars 2-6 cells wide at Poisson-disk spacing 7 (the configurations'
rule; spacing 7 rather than 8 reaches their ~30% fill with 2-6 wide pillars),
one-cell floor slabs every 8-16 cells with 20% 2x2 holes (seed 1); per-cell
log-stiffness theta ~ N(0, 0.5^2) (e = exp(theta), g = 0.4 e); gravity of weight
1 per free cell (a sixth on each face, -z) plus 64 point loads of magnitude
U(1, 10) in uniformly random directions on random free nodes. Synthetic,
generated here.

One step = classify (joined to held -> gone) + mixed-precision solve to 1e-8 that
frees its multigrid hierarchy when done + the adjoint of the soft-max (beta = 20)
bond stress w.r.t. per-cell log-stiffness (a second solve, mixed, 1e-8, which
recomputes the hierarchy instead of reading a stored one, and frees it again).
The timed steps run with the loads and the solution resident in device memory.
Alongside: the same step in plain FP64 (like for like), the forward solves
alone, the error of each result against a tight FP64 solve, and the CPU
baseline.
"""
import argparse
import json
import re
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "parallax-world_amd"))

import statics_hip as sh  # noqa: E402

PEAK_HBM_GBS = 8000.0
COPY_BYTES = 1 << 30
# The roofline kernel is the CG's K p with the fused p.Ap in FP64 (epilogue EpiDot<double>),
# in whichever schedule the autotuner chose: k_apply (node-centric, one dispatch per
# application) or k_apply_cells (cell-centric, one dispatch per brick class).
ROOFLINE_EPILOGUE = "EpiDot<double>"


def traffic_file(kernel):
    return os.path.join(HERE, "profiles", f"{kernel}_epidot_f64_256_traffic.txt")


def _pmc_per_dispatch(csv_path, counter):
    """Mean counter value per dispatch for the roofline kernel and for the three calibration
    copies (the last three copy dispatches of the run), and the roofline kernel's name."""
    import csv
    apply, copies, names = [], [], set()
    with open(csv_path) as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter:
                continue
            name = row["Kernel_Name"]
            v = float(row["Counter_Value"])
            if ROOFLINE_EPILOGUE in name:
                apply.append(v)
                m = re.search(r"(k_apply\w*)<", name)
                names.add(m.group(1) if m else "k_apply")
            elif "copyBuffer" in name:
                copies.append((int(row["Dispatch_Id"]), v))
    if len(names) > 1:
        raise RuntimeError(f"{csv_path}: the run used more than one schedule for the roofline kernel: {names}")
    copies = [v for _, v in sorted(copies)[-3:]]
    vals = {"apply": apply, "copy": copies}
    mean = {k: (sum(v) / len(v) if v else None) for k, v in vals.items()}
    return mean, {k: len(v) for k, v in vals.items()}, (names.pop() if names else None)


def copy_calibration():
    """Three device-to-device copies of COPY_BYTES through the HIP runtime (its copy kernel
    reads and writes exactly that many bytes): the reference for the traffic counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(COPY_BYTES)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(COPY_BYTES))
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(COPY_BYTES), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def traffic_summary(fetch_csv, write_csv, schedule):
    """FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of the roofline kernel, each scaled by the
    ratio of a 1 GiB device copy's known bytes to what the same counter reads for it, and
    summed over the dispatches of one application of the operator."""
    fe, nf, kname = _pmc_per_dispatch(fetch_csv, "FETCH_SIZE")
    wr, nw, kname_w = _pmc_per_dispatch(write_csv, "WRITE_SIZE")
    if kname is None or kname != kname_w:
        raise RuntimeError(f"roofline kernel not found in both runs ({kname}, {kname_w})")
    per_app = 2 if kname == "k_apply_cells" else 1
    kf = COPY_BYTES / (fe["copy"] * 1024.0)
    kw = COPY_BYTES / (wr["copy"] * 1024.0)
    fetch_b = fe["apply"] * 1024.0 * kf * per_app
    write_b = wr["apply"] * 1024.0 * kw * per_app
    lines = [
        f"# HBM traffic per application of {kname}<double, EpiDot<double>> (the CG's K p with the fused p.Ap),",
        f"# 256^3 bench workload; {per_app} dispatch(es) per application",
        "# commands (GPU, from the repository root, cd /tmp && export TMPDIR=/tmp first):",
        f"#   STATICS_HIP_SCHEDULE={schedule} rocprofv3 --pmc FETCH_SIZE --kernel-include-regex "
        "'EpiDot<double>|copyBuffer' -d <dir> -o fetch --output-format csv -- python3 bench.py --steps 1 "
        "--warmup 0 --cpu-sample 0 --no-fp64 --no-errors --copy-calib",
        "#   the same with --pmc WRITE_SIZE and -o write",
        f"#   python3 bench.py --traffic <fetch_counter_collection.csv> <write_counter_collection.csv> "
        f"--traffic-schedule {schedule}",
        f"kernel: {kname}",
        f"dispatches_per_application: {per_app}",
        f"fetch_kib_per_dispatch_raw: {fe['apply']:.1f} (dispatches {nf['apply']})",
        f"write_kib_per_dispatch_raw: {wr['apply']:.1f} (dispatches {nw['apply']})",
        f"copy_1GiB_fetch_kib_raw: {fe['copy']:.1f} (dispatches {nf['copy']}) -> fetch scale {kf:.4f}",
        f"copy_1GiB_write_kib_raw: {wr['copy']:.1f} (dispatches {nw['copy']}) -> write scale {kw:.4f}",
        f"fetch_bytes_per_launch: {fetch_b:.6e}",
        f"write_bytes_per_launch: {write_b:.6e}",
        f"traffic_bytes_per_launch: {fetch_b + write_b:.6e}",
    ]
    path = traffic_file(kname)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def measured_traffic(kernel):
    """HBM bytes per application of the roofline kernel from the committed counter summary
    for that kernel, or None."""
    try:
        with open(traffic_file(kernel)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        return None
    return None


def structure(n, seed, sigma=0.5, spacing=7.0, wmin=2, wmax=6):
    occ, theta = sh.generate(n, n, n, seed=seed, sigma=sigma, pillar_spacing=spacing, pillar_min=wmin,
                             pillar_max=wmax)
    return occ, theta


def node_loads(occ, faces, joined, held, rng, n_points=64):
    """Gravity (weight 1 per free cell, a sixth per face, -z) and n_points point loads."""
    o = faces[:, 0]
    plus = faces[:, 1:4]
    minus = plus - np.eye(3, dtype=np.int32)[o]
    n = occ.shape[0]

    def free(c):
        ok = np.all((c >= 0) & (c < n), axis=1)
        cc = np.clip(c, 0, n - 1)
        return ok & (occ[cc[:, 0], cc[:, 1], cc[:, 2]] == 1)

    cnt = free(plus).astype(np.float64) + free(minus)
    f = np.zeros((len(faces), 6))
    f[:, 2] = -cnt / 6.0
    cand = np.flatnonzero(joined.astype(bool) & ~held.astype(bool))
    pick = rng.choice(cand, size=min(n_points, len(cand)), replace=False)
    d = rng.normal(size=(len(pick), 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    f[pick, :3] += d * rng.uniform(1.0, 10.0, size=(len(pick), 1))
    return f


def cpu_baseline(size, threads, seed=1, points=64, max_iter=100000, progress=0, spacing=7.0):
    """The oracle (faithful restatement, FP64, assembled blocks, smoothed aggregation) on a size^3
    sample of the same generator and loads (gravity + point loads). progress > 0: the oracle
    prints a line per level built and every `progress` CG iterations (stderr)."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as orc
    orc.lib().orc_set_progress(progress)
    occ, theta = structure(size, seed, spacing=spacing)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    mats = np.zeros((len(sites), 6))
    mats[:, 0] = 1.0
    mats[:, 1] = e
    mats[:, 2] = g
    mats[:, 3:] = 1.0
    orc.lib().orc_set_threads(threads)
    say = (lambda m: print(f"[cpu] {m}", file=sys.stderr, flush=True)) if progress > 0 else (lambda m: None)
    say(f"{size}^3: {len(sites)} cells, building the lattice")
    t0 = time.perf_counter()
    L = orc.Lattice(0, sites, mats, held)
    t_build = time.perf_counter() - t0
    say(f"lattice: {L.n_nodes} nodes, {L.n_bonds} bonds ({t_build:.1f} s)")
    t1 = time.perf_counter()
    joined = L.joined_to_held()
    gone = (joined == 0).astype(np.uint8)
    f = np.zeros((L.n_nodes, 6))
    f[:, 2] = -L.m / (orc.UNIT ** 3)  # weight 1 per free cell, a sixth per face
    rng = np.random.default_rng(seed)
    free = np.flatnonzero((L.held == 0) & (gone == 0))
    pick = rng.choice(free, size=min(points, len(free)), replace=False)
    d = rng.normal(size=(len(pick), 3))
    f[pick, :3] += d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(1.0, 10.0, size=(len(pick), 1))
    x, st = L.motions(f, gone=gone, tol=1e-8, max_iter=max_iter)
    t_solve = time.perf_counter() - t1
    dof = 6 * int(((L.held == 0) & (gone == 0)).sum())
    # the true relative residual reached, ||f - K x|| / ||f|| (on samples small enough that assembling K
    # once more is cheap; the CG itself stops on its recursive residual, bodies.rs:350-352)
    rel = None
    if dof < 6_000_000:
        part = (L.held == 0) & (gone == 0)
        rel = float(np.linalg.norm(f[part] - L.apply(x, gone=gone)[part]) / np.linalg.norm(f[part]))
    return dict(seconds=t_solve, build_s=t_build, iters=st["iters"], dof=dof, levels=st["levels"],
                setup_s=st["setup_s"], solve_s=st["solve_s"], capped=bool(st["capped"]), rel_residual=rel,
                cpu=orc.host_cpu(), threads=threads)


def short_solve(args):
    """The --size structure uploaded, every operator schedule timed (the autotuner) and a mixed solve of
    --short-solve iterations under the bench's loads: a brief run of every operator kernel, for counters."""
    import ctypes
    occ, theta = structure(args.size, args.seed, spacing=args.spacing)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    ctx = sh.Context().upload(sites, held, e, g)
    joined = ctx.joined_to_held()
    loads = node_loads(occ, ctx.node_faces(), joined, ctx.node_held(), np.random.default_rng(args.seed),
                       n_points=args.points)
    ctx._check(ctx.L.st_set_loads(ctx.h, sh._p(np.ascontiguousarray(loads))), "st_set_loads")
    c = ctypes.c_int64()
    ctx._check(ctx.L.st_classify(ctx.h, ctypes.byref(c)), "st_classify")
    o = sh.StSolveOpts(sh.MIXED, 1e-8, args.short_solve, 0, 0, 1)
    s = sh.StStats()
    rc = ctx.L.st_solve_resident(ctx.h, ctypes.byref(o), ctypes.byref(s))
    print(f"[bench] short solve: rc {rc}, {s.iters} iterations, setup {s.setup_ms:.1f} ms", file=sys.stderr,
          flush=True)


def cpu_full_record(edge, args):
    """The committed record of a --cpu-full run on the same workload (edge, spacing, seed, point
    loads): the JSON line in profiles/cpu_oracle_<n>.txt whose edge is `edge`, or None."""
    path = os.path.join(HERE, "profiles", f"cpu_oracle_{args.size}.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("{"):
                    r = json.loads(line)
                    if (r.get("edge"), r.get("spacing"), r.get("seed"), r.get("points")) == (
                            edge, args.spacing, args.seed, args.points):
                        return r
    except OSError:
        pass
    return None


def cpu_full(args):
    """The CPU oracle on the whole workload (not a bounded sample): classify + motions to 1e-8 at
    --cpu-full edge, with progress lines, at most --cpu-max-iter iterations; one JSON line."""
    import resource
    t = time.perf_counter()
    cb = cpu_baseline(args.cpu_full, args.cpu_threads, args.seed, args.points, max_iter=args.cpu_max_iter,
                      progress=10, spacing=args.spacing)
    cb.update(edge=args.cpu_full, spacing=args.spacing, seed=args.seed, points=args.points,
              wall_s=time.perf_counter() - t, nproc=os.cpu_count(),
              maxrss_gb=resource.getrusage(resource.RUSAGE_SELF).ru_maxrss / 1e6)
    print(json.dumps(cb), flush=True)


def batch_structures(count, seed):
    """The batch workload: side L = 16 * 8^U (16..128), solid fraction about U(0.15, 0.40) reached
    through the pillar spacing (widths 2-6; fill ~0.32 at spacing 6.5, ~ 1/spacing^2), lognormal stiffness
    sigma 0.5."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        side = int(16 * 8 ** rng.uniform(0.0, 1.0))
        phi = rng.uniform(0.15, 0.40)
        spacing = 6.5 * np.sqrt(0.32 / phi)
        occ, theta = sh.generate(side, side, side, seed=seed * 100003 + i, sigma=0.5, pillar_spacing=spacing,
                                 pillar_min=2, pillar_max=6)
        out.append(sh.cells_from_occ(occ, theta))
    return out


def batch_bench(args):
    """Solves the batch in st_solve_batch calls of about --batch-cells free cells each (mixed, 1e-8,
    gravity), after one untimed call on the first chunk; reports structures solved per second."""
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    t = time.perf_counter()
    sets = batch_structures(args.batch, args.seed)
    cells = [int((h == 0).sum()) for _, h, _, _ in sets]
    sides = [int(s[0].max()) + 1 for s in sets]
    log(f"[bench] batch of {len(sets)}: {sum(cells)} free cells, sides {min(sides)}..{max(sides)} "
        f"({time.perf_counter() - t:.1f}s)")
    chunks, cur, acc = [], [], 0
    for i, c in enumerate(cells):
        if cur and acc + c > args.batch_cells:
            chunks.append(cur)
            cur, acc = [], 0
        cur.append(i)
        acc += c
    if cur:
        chunks.append(cur)
    ctx = sh.Context()
    ctx.solve_batch([sets[i] for i in chunks[0]], precision=sh.MIXED, tol=1e-8, max_iter=20000, want_x=False)
    t0 = time.perf_counter()
    stats, host_s, device_s, mem = [], 0.0, 0.0, 0
    for k, ch in enumerate(chunks):
        tc = time.perf_counter()
        _, st = ctx.solve_batch([sets[i] for i in ch], precision=sh.MIXED, tol=1e-8, max_iter=20000, want_x=False)
        wall = time.perf_counter() - tc
        dev = (st[0].setup_ms + st[0].solve_ms) * 1e-3
        host_s += wall - dev
        device_s += dev
        mem = max(mem, max(s.mem_peak for s in st))
        stats += st
        log(f"[bench] chunk {k}: {len(ch)} structures, {sum(cells[i] for i in ch)} cells, {wall * 1e3:.0f} ms "
            f"(host brick build, upload and classification {(wall - dev) * 1e3:.0f} ms), iterations max "
            f"{max(s.iters for s in st)}, setup {st[0].setup_ms:.0f} ms, solve {st[0].solve_ms:.0f} ms, device memory "
            f"peak {max(s.mem_peak for s in st) / 1e9:.1f} GB")
    elapsed = time.perf_counter() - t0
    bad = [i for i, s in enumerate(stats) if s.status != 0]
    out = {
        "metric": "structures solved/sec to 1e-8 residual, batch of uneven structures (mixed precision)",
        "value": len(sets) / elapsed, "unit": "structures/sec", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": elapsed * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 CG + f32 V-cycle", "data": "synthetic",
        "config": {"workload": f"batch of {len(sets)} pillar/slab structures, side 16*8^U, fill ~U(0.15,0.40), "
                               f"lognormal sigma 0.5, gravity; {len(chunks)} st_solve_batch call(s) of <= "
                               f"{args.batch_cells:.0f} free cells", "structures": len(sets),
                   "free_cells": sum(cells), "chunks": len(chunks)},
        "iterations_max": max(s.iters for s in stats), "iterations_mean": float(np.mean([s.iters for s in stats])),
        "failed": bad,
        "host_s": host_s, "device_s": device_s, "device_memory_peak_gb": mem / 1e9,
        "iteration_graphs": "each chunk of 8 PCG iterations is one captured-graph launch over every structure",
    }
    print(json.dumps(out), flush=True)


def fd_check(args, eps=1e-4, beta=20.0):
    """Adjoint gradients against central differences of FP64 solves (tol 1e-10: the FP64 floor of
    these systems lies near 1e-11; the achieved residuals are reported) at random free cells."""
    import ctypes
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    n = args.size
    occ, theta = structure(n, args.seed)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    ctx = sh.Context().upload(sites, held, e, g)
    joined = ctx.joined_to_held()
    nheld = ctx.node_held()
    faces = ctx.node_faces()
    f = node_loads(occ, faces, joined, nheld, np.random.default_rng(args.seed), n_points=0)
    t0 = time.perf_counter()
    x0, st0 = ctx.solve(f, precision=sh.FP64, tol=1e-10, max_iter=20000, check=False)
    worst = st0.rel_residual
    y0 = ctx.apply(x0)
    grads = {}
    for name, obj in (("compliance", sh.COMPLIANCE), ("softmax_stress", sh.SOFTMAX_STRESS)):
        o = sh.StSolveOpts(sh.FP64, 1e-10, 20000, 0, 0)
        Jc, d, sa = ctypes.c_double(), np.zeros(len(sites)), sh.StStats()
        ctx.L.st_backward(ctx.h, obj, beta, ctypes.byref(o), ctypes.byref(Jc), sh._p(d), ctypes.byref(sa))
        worst = max(worst, sa.rel_residual)
        grads[name] = (Jc.value, d)
    t_adj = time.perf_counter() - t0
    log(f"[bench] adjoints done ({t_adj:.1f} s), worst residual {worst:.1e}")
    rng = np.random.default_rng(args.seed + 1)
    free = np.flatnonzero(held == 0)
    pick = rng.choice(free, size=min(args.fd_check, len(free)), replace=False)
    errs = {k: [] for k in grads}
    absd = {k: [] for k in grads}
    t1 = time.perf_counter()
    for i, c in enumerate(pick):
        Jpm = {k: [] for k in grads}
        for sgn in (1.0, -1.0):
            e2, g2 = e.copy(), g.copy()
            e2[c] *= np.exp(sgn * eps)
            g2[c] *= np.exp(sgn * eps)
            ctx.set_stiffness(e2, g2)
            # x(theta +- eps) = x0 + d with K(theta +- eps) d = -(K(theta +- eps) - K0) x0: solving for
            # the difference d puts the solve's error on d, not on x, so C(+) - C(-) = f.d(+) - f.d(-)
            # keeps its digits (two solves of x to 1e-10 would leave ~1e-10 C of noise in each C)
            b = y0 - ctx.apply(x0)
            d, std = ctx.solve(b, precision=sh.FP64, tol=1e-10, max_iter=20000, check=False)
            # the context's solution becomes x0 + d (already converged) for the soft-max objective
            xp, stp = ctx.solve(f, precision=sh.FP64, tol=1e-10, max_iter=20000, x0=x0 + d, check=False)
            worst = max(worst, stp.rel_residual, std.rel_residual * float(np.abs(b).max() > 0))
            Jpm["compliance"].append(float((f * d).sum()))
            Jpm["softmax_stress"].append(ctx.objective(sh.SOFTMAX_STRESS, beta))
        for k in grads:
            fd = (Jpm[k][0] - Jpm[k][1]) / (2 * eps)
            dJ = grads[k][1]
            scale = max(abs(fd), 1e-3 * np.abs(dJ).max())
            errs[k].append(abs(dJ[c] - fd) / scale)
            absd[k].append((abs(dJ[c] - fd), abs(fd)))
        if True:
            log(f"[bench] fd cell {i + 1}/{len(pick)}: max rel err compliance {max(errs['compliance']):.2e}, "
                f"soft-max {max(errs['softmax_stress']):.2e}")
    ctx.set_stiffness(e, g)
    out = {
        "check": f"adjoint gradients vs central differences, {n}^3 structure, {len(pick)} random free cells, "
                 f"eps {eps}, FP64 solves to 1e-10, beta {beta}; C(+-eps) - C(0) = f.d with d solved from "
                 f"K(+-eps) d = -(K(+-eps) - K0) x0",
        "worst_rel_residual": worst,
        "free_cells": int(len(free)),
        "max_rel_err_compliance": float(max(errs["compliance"])),
        "max_rel_err_softmax_stress": float(max(errs["softmax_stress"])),
        "per_cell_rel_err": "|dJ_c - fd_c| / max(|fd_c|, 1e-3 max_c |dJ_c|)",
        "norm_rel_err_compliance": max(a for a, _ in absd["compliance"]) / max(b for _, b in absd["compliance"]),
        "norm_rel_err_softmax_stress": max(a for a, _ in absd["softmax_stress"]) / max(b for _, b in absd["softmax_stress"]),
        "norm_rel_err": "max_c |dJ_c - fd_c| / max_c |fd_c| over the sampled cells",
        "bar": 1e-4,
        "invariant_sum_dC_plus_C": float(grads["compliance"][1].sum() + grads["compliance"][0]) / grads["compliance"][0],
        "invariant_sum_dJ_softmax": float(grads["softmax_stress"][1].sum() / np.abs(grads["softmax_stress"][1]).sum()),
        "adjoint_s": t_adj, "fd_s": time.perf_counter() - t1,
    }
    print(json.dumps(out), flush=True)


def settle_cpu_sample(size, seed, quantile, threads, spacing=7.0):
    """The oracle's settle (bodies.rs:511-541, gravity -y, tol 1e-6) on a size^3 structure of the
    same generator, strengths at the same quantile of its standing used shares."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as orc
    occ, theta = structure(size, seed, spacing=spacing)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    mats = np.zeros((len(sites), 6))
    mats[:, 0], mats[:, 1], mats[:, 2], mats[:, 3:] = 1.0, e, g, 1.0
    orc.lib().orc_set_threads(threads)
    u = orc.Lattice(0, sites, mats, held).used_standing(tol=1e-6)
    mats[:, 3:] = float(np.quantile(u[u > 0], quantile))
    L = orc.Lattice(0, sites, mats, held)
    t0 = time.perf_counter()
    o = L.settle(tol=1e-6)
    sec = time.perf_counter() - t0
    dof = 6 * int((L.held == 0).sum())
    return dict(seconds=sec, rounds=o["rounds"], fallen=o["fallen"], dof=dof)


def settle_bench(args):
    """bodies::settle on the device (st_settle) for one --settle^3 structure of the bench's generator
    under gravity along -y (the reference's forces), every strength at the --settle-quantile of the
    standing structure's used shares, so that the most used bonds break and breaking cascades over
    rounds; mixed solves to the reference's settle tolerance 1e-6 under its stopping rule (the
    recursive residual decides), rounds warm-started."""
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    n = args.settle
    occ, theta = structure(n, args.seed, spacing=args.spacing)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    ctx = sh.Context().upload(sites, held, e, g)
    faces = ctx.node_faces()
    joined = ctx.joined_to_held()
    nheld = ctx.node_held()
    f = node_loads(occ, faces, joined, nheld, np.random.default_rng(args.seed), n_points=0)
    f = np.ascontiguousarray(f[:, [0, 2, 1, 3, 4, 5]])  # the weights along -y
    x, st = ctx.solve(f, precision=sh.MIXED, tol=1e-6, max_iter=20000)
    u = ctx.bond_used(np.ones((len(sites), 3)), x)
    s = float(np.quantile(u[u > 0], args.settle_quantile))
    del u
    strength = np.full((len(sites), 3), s)
    log(f"[bench] settle {n}^3: {ctx.cells} free cells, strength {s:.4g} (quantile {args.settle_quantile})")
    times, last = [], None
    for i in range(args.warmup + args.steps):
        ctx.set_beams(None)  # every beam alive again
        t0 = time.perf_counter()
        gone, beams, rounds, st = ctx.settle(strength, f, precision=sh.MIXED, tol=1e-6, max_iter=20000,
                                             warm_start=True, refresh=-1)
        dt = time.perf_counter() - t0
        log(f"[bench] settle {i}: {dt * 1e3:.1f} ms, {rounds} rounds, {st.iters} CG iterations, "
            f"{int(gone.sum())} nodes gone")
        if i >= args.warmup:
            times.append(dt)
        last = (gone, beams, rounds, st)
    gone, beams, rounds, st = last
    ms = 1e3 * sum(times) / len(times)
    dof = 6 * int(st.dof_nodes) if st.dof_nodes else 6 * ctx.nodes
    dof0 = 6 * int((nheld == 0).sum())
    cpu = None
    if args.settle_cpu_sample > 0:
        cb = settle_cpu_sample(args.settle_cpu_sample, args.seed, args.settle_quantile, args.cpu_threads,
                               spacing=args.spacing)
        per_s = (1.0 / cb["seconds"]) * (cb["dof"] * cb["rounds"]) / (dof0 * rounds)
        cpu = {"value": per_s, "unit": f"settles/sec ({n}^3, extrapolated)", "cores": args.cpu_threads,
               "kind": "port",
               "sample": f"oracle settle (C++ FP64 restatement of statics/src/bodies.rs:511-541) on the "
                         f"{args.settle_cpu_sample}^3 structure of the same generator and strength rule: "
                         f"{cb['seconds']:.2f} s, {cb['rounds']} rounds, {cb['dof']} DOF; scaled by DOF x rounds "
                         f"to the {n}^3 settle's {dof0} DOF x {rounds} rounds"}
    out = {
        "metric": f"settles/sec of one {n}^3 structure (bodies::settle on the device)",
        "value": 1e3 / ms, "unit": "settles/sec", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 CG + f32 V-cycle", "data": "synthetic",
        "config": {"workload": f"one {n}^3 pillar/slab structure (pillar spacing {args.spacing:g}), lognormal "
                               f"stiffness sigma 0.5, gravity along -y, strength at the {args.settle_quantile} "
                               f"quantile of the standing used shares; step = st_settle (classify, solve to 1e-6, "
                               f"break, until nothing breaks)",
                   "lattice": n, "free_cells": int(ctx.cells), "dof_standing": dof0},
        "rounds": rounds, "cg_iterations": int(st.iters), "nodes_gone": int(gone.sum()),
        "beams_broken": int(np.unpackbits((beams[held == 0] ^ np.uint16(0x7FFF)).view(np.uint8)).sum()),
        "setup_ms_total": st.setup_ms, "solve_ms_total": st.solve_ms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def grid_of(occ):
    """The occupied sites of occ (x, y, z) as a lattice::Grid (lattice/src/lib.rs:182-205): 8^3
    places z-major, each non-empty one its own 512 sites, z-major within; Grid.solid bits."""
    n = occ.shape
    nb = [(v + 7) // 8 for v in n]
    occ_p = np.zeros([8 * v for v in nb], np.uint8)
    occ_p[:n[0], :n[1], :n[2]] = occ > 0
    blocks = occ_p.reshape(nb[0], 8, nb[1], 8, nb[2], 8).transpose(4, 2, 0, 5, 3, 1).reshape(-1, 512)
    nonempty = blocks.any(axis=1)
    bricks = np.where(nonempty, np.cumsum(nonempty) - 1, -1).astype(np.int32)
    bits = blocks[nonempty].reshape(-1)
    solid = np.packbits(bits, bitorder="little").view(np.uint32)
    return bricks, solid, len(bits)


def ingest_bench(args):
    """solids + hold_bulk from a lattice::Grid on the device (st_grid_cells, the count call and
    the fill call) for the --ingest^3 structure of the bench's generator; the oracle's
    hold_bulk on a 128^3 sample as the CPU baseline (scaled by cells)."""
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    n, radius = args.ingest, 0.3
    occ, _ = structure(n, args.seed, spacing=args.spacing)
    bricks, solid, nsites = grid_of(occ)
    del occ
    ctx = sh.Context()
    times = []
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        sites, held, index = ctx.grid_cells((0, 0, 0), (n, n, n), bricks, solid, nsites, 0, radius)
        dt = time.perf_counter() - t0
        if i >= args.warmup:
            times.append(dt)
        log(f"[bench] ingest {i}: {dt * 1e3:.1f} ms, {len(sites)} cells, {int(held.sum())} held")
    ms = 1e3 * sum(times) / len(times)
    cpu = None
    if args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as orc
        m = min(args.cpu_sample * 2, n)
        occ_s, _ = structure(m, args.seed, spacing=args.spacing)
        sx, sy, sz = np.nonzero(occ_s > 0)
        s_sites = np.stack([sx, sy, sz], axis=1).astype(np.int64)
        t0 = time.perf_counter()
        orc.hold_bulk(0, s_sites, radius)
        sec = time.perf_counter() - t0
        cpu = {"value": len(s_sites) / sec, "unit": "cells/sec", "cores": 1, "kind": "port",
               "sample": f"oracle hold_bulk (statics/src/lattice.rs:107-119 restated, hash set) on the {m}^3 "
                         f"structure's {len(s_sites)} cells, radius {radius}: {sec:.2f} s (the cells are given, "
                         f"no grid decode)"}
    out = {"metric": f"cells/sec classified from a lattice::Grid ({n}^3 tile: solids + hold_bulk on the device)",
           "value": len(sites) / (ms * 1e-3), "unit": "cells/sec", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u32 bits", "data": "synthetic",
           "config": {"workload": f"the bench generator's {n}^3 structure as a lattice::Grid, hold radius {radius} m; "
                                  f"step = st_grid_cells count + fill, host copies included",
                      "sites": n ** 3, "cells": int(len(sites)), "held": int(held.sum())},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


FACES6 = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]], np.int64)


def level0_lattice(sites, held, mat):
    """lattice::build (statics/src/lattice.rs:124-163) at level 0, all cells of one material, in
    numpy: nodes in the order build makes them (each free cell's faces in VOXEL_FACES order), its
    beams cell by cell, PER_BEAM bonds each (lattice.rs:76-84)."""
    s = sh.UNIT
    sites = np.asarray(sites, np.int64)
    held = np.asarray(held, bool)
    lo_all = sites.min(0) - 1
    dims = sites.max(0) - lo_all + 2
    hgrid = np.zeros(dims, bool)
    hs = sites[held] - lo_all
    hgrid[hs[:, 0], hs[:, 1], hs[:, 2]] = True
    free = sites[~held]
    other = free[:, None, :] + FACES6[None]
    plus = (np.arange(6) % 2 == 1)[None, :, None]
    lo = np.where(plus, free[:, None, :], other)
    hi = np.where(plus, other, free[:, None, :])
    rel = lo - lo_all
    code = ((rel[..., 0] * dims[1] + rel[..., 1]) * dims[2] + rel[..., 2]) * 3 + np.arange(6)[None] // 2
    uniq, first, inv = np.unique(code.ravel(), return_index=True, return_inverse=True)
    rank = np.empty(len(uniq), np.int64)
    rank[np.argsort(first)] = np.arange(len(uniq))
    node = rank[inv].reshape(-1, 6)
    nn = len(uniq)
    centre = lambda q: (q.astype(np.float64) + 0.5) * s  # noqa: E731
    lat = sh.GatheredLattice(0, 0, 0)
    lat.level = 0
    x = np.zeros((nn, 3))
    x[node.ravel()] = ((centre(free[:, None, :].repeat(6, 1)) + centre(other)) * 0.5).reshape(-1, 3)
    lat.x = x
    lat.m = np.bincount(node.ravel(), minlength=nn) * (mat[0] * s * s * s / 6.0)
    lat.node_mat = np.tile(np.asarray(mat, np.float64), (nn, 1))
    oh = other - lo_all
    nheld = np.zeros(nn, np.uint8)
    nheld[node.ravel()] = hgrid[oh[..., 0], oh[..., 1], oh[..., 2]].ravel()
    lat.held = nheld
    around = np.zeros((nn, 2, 3), np.int64)
    around[node.ravel(), 0] = lo.reshape(-1, 3)
    around[node.ravel(), 1] = hi.reshape(-1, 3)
    lat.around = around
    pa, pb = np.triu_indices(6, 1)
    ea, eb = node[:, pa], node[:, pb]
    keep = ~(nheld[ea].astype(bool) & nheld[eb].astype(bool))
    ea, eb = ea[keep], eb[keep]
    xa, xb = x[ea], x[eb]
    unit = lambda v: v / np.linalg.norm(v, axis=-1, keepdims=True)  # noqa: E731
    nv = unit(xb - xa)
    ref = np.where((np.abs(nv[:, 0]) < 0.9)[:, None], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0])
    e1 = unit(np.cross(nv, ref))
    e2 = np.cross(nv, e1)
    area = 0.336 * s * s
    r = np.sqrt(area / (2.0 * np.pi))
    mid = (xa + xb) * 0.5
    off = np.stack([e1 * r, -e1 * r, e2 * r, -e2 * r], 1)
    nb = len(ea) * 4
    lat.ij = np.stack([ea, eb], 1).repeat(4, 0)
    lat.p = (mid[:, None, :] + off).reshape(-1, 3)
    lat.n = nv.repeat(4, 0)
    lat.a = np.full(nb, area / 4.0)
    lat.d = np.linalg.norm(xb - xa, axis=1).repeat(4)
    lat.bond_mat = np.tile(np.asarray(mat, np.float64), (nb, 1))
    return lat


def gather_bench(args):
    """gather::gathered on the device (st_gather: the host's bookkeeping, the cells measured on the
    device, the coarse lattice assembled) for the bench generator's --gather^3 structure built as a
    level-0 lattice of stone; the oracle's gathered on a smaller structure as the CPU baseline."""
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    n = args.gather
    stone = (2500.0, 30e9, 12e9, 5e6, 100e6, 10e6)  # statics/src/model.rs:53
    occ, _ = structure(n, args.seed, spacing=args.spacing)
    sites, held, _, _ = sh.cells_from_occ(occ)
    del occ
    t0 = time.perf_counter()
    lat = level0_lattice(sites, held, stone)
    log(f"[bench] gather: {len(sites)} cells ({int(held.sum())} held), {len(lat.m)} nodes, {len(lat.a)} bonds "
        f"built in {time.perf_counter() - t0:.1f} s")
    ctx = sh.Context()
    times = []
    for i in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        g = ctx.gather(lat)
        dt = time.perf_counter() - t0
        if i >= args.warmup:
            times.append(dt)
        log(f"[bench] gather {i}: {dt * 1e3:.1f} ms, {len(g.m)} nodes, {len(g.a)} bonds, "
            f"weight kept to {abs(g.m.sum() / lat.m.sum() - 1):.1e}")
    ms = 1e3 * sum(times) / len(times)
    cpu = None
    if args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as orc
        m = min(args.gather_cpu_sample, n)
        occ_s, _ = structure(m, args.seed, spacing=args.spacing)
        s_sites, s_held, _, _ = sh.cells_from_occ(occ_s)
        olat = orc.Lattice(0, s_sites.astype(np.int64), stone, s_held)
        orc.lib().orc_set_threads(args.cpu_threads)
        t0 = time.perf_counter()
        olat.gathered()
        sec = time.perf_counter() - t0
        cpu = {"value": len(s_sites) / sec, "unit": "cells/sec", "cores": args.cpu_threads, "kind": "port",
               "sample": f"oracle gathered (statics/src/gather.rs:228-395 restated, cells over {args.cpu_threads} "
                         f"threads as the reference's `each`) on the {m}^3 structure's {len(s_sites)} cells: "
                         f"{sec:.2f} s"}
    out = {"metric": f"level-0 cells/sec gathered to level 1 ({n}^3 structure, gather::gathered on the device)",
           "value": len(sites) / (ms * 1e-3), "unit": "cells/sec", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"the bench generator's {n}^3 structure, all stone, as a level-0 lattice; "
                                  f"step = st_gather + st_gather_read, host bookkeeping and copies included",
                      "cells": int(len(sites)), "nodes": int(len(lat.m)), "bonds": int(len(lat.a)),
                      "coarse_nodes": int(len(g.m)), "coarse_bonds": int(len(g.a))},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def mechanics_system(n, seed):
    """The mechanics system of the bench generator's n^3 structure (mech_systems.system: the
    engine's assembly, mechanics/src/lib.rs:638-676): every solid cell a wood cube of 0.125 m,
    turned up to 0.2 rad, the corners of the held cells held; dt 1/60."""
    import mech_systems as ms
    occ, _ = structure(n, seed)
    sites = np.argwhere(occ > 0)
    hc = np.argwhere(occ == 2)
    side = n + 2
    code = lambda k: (k[:, 0] * side + k[:, 1]) * side + k[:, 2]  # noqa: E731
    held_codes = np.unique(code((hc[:, None, :] + ms.CORNERS[None]).reshape(-1, 3)))
    held = lambda k: np.isin(code(k), held_codes)  # noqa: E731
    s, b, _, _ = ms.system(sites, held, h=sh.UNIT, turn=0.2, seed=seed)
    return s, b, len(sites)


def mechanics_cpu_sample(n, seed):
    """The oracle's Solver::new + solve to 1e-3 on the n^3 system (one thread, as the reference's)."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as orc
    s, b, cells = mechanics_system(n, seed)
    o_sys = orc.MechSystem(s.sites, s.at, s.diag, s.ij, s.blocks)
    v0 = np.random.default_rng(seed + 1).normal(size=b.shape) * 1e-2  # as mechanics_bench starts
    orc.lib()  # built (for this host's CPU) before the clock starts
    t0 = time.perf_counter()
    o = orc.MechSolver(o_sys, sh.UNIT)
    _, rounds = o.solve(o_sys, b, x0=v0, tol=1e-3)
    return {"seconds": time.perf_counter() - t0, "unknowns": s.n, "rounds": rounds, "cells": cells}


def mechanics_bench(args):
    """mechanics::solve::Solver on the device for the system of one --mechanics^3 structure of
    the bench generator: a step is what the engine does when its kept levels do not fit
    (mechanics/src/lib.rs:741-746): Solver::new, then solve to the engine's 1e-3 from the
    current speeds. Reported beside it: the solve alone, and a solve to 1e-8."""
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    n = args.mechanics
    t = time.perf_counter()
    s, b, cells = mechanics_system(n, args.seed)
    log(f"[bench] mechanics {n}^3: {cells} cells, {s.n} unknowns, {len(s.ij)} pairs "
        f"(built in {time.perf_counter() - t:.1f}s)")
    v0 = np.random.default_rng(args.seed + 1).normal(size=b.shape) * 1e-2
    times, solve_ms, rounds, solver = [], [], None, None
    for i in range(args.warmup + args.steps):
        if solver is not None:
            solver.close()
        t0 = time.perf_counter()
        solver = sh.MechSolver(s, sh.UNIT)
        t1 = time.perf_counter()
        if i >= args.warmup:
            solver.kernel_timing(1)
        x, rounds = solver.solve(s, b, x0=v0, tol=1e-3)
        t2 = time.perf_counter()
        log(f"[bench] mechanics step {i}: new {1e3 * (t1 - t0):.1f} ms, solve {1e3 * (t2 - t1):.1f} ms, "
            f"{rounds} rounds")
        if i >= args.warmup:
            times.append(t2 - t0)
            solve_ms.append(1e3 * (t2 - t1))
            launches, kp_ms, kp_bytes = solver.kernel_timing(-1)
            solver.kernel_timing(0)
    info = solver.info()
    ms_step = 1e3 * sum(times) / len(times)
    t0 = time.perf_counter()
    _, rounds_tight = solver.solve(s, b, x0=v0, tol=1e-8)
    tight_ms = 1e3 * (time.perf_counter() - t0)
    avg = kp_ms / max(launches, 1)
    roof = {"bound": "hbm", "kernel": "k_spmv<3, 0> (the CG's K p on the three-unknown level: a row of 3x3 blocks "
                                      "per unknown, both halves of each pair)",
            "achieved": kp_bytes / (avg * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
            "frac": kp_bytes / (avg * 1e-3) / 1e9 / 8000.0, "traffic": None, "launches": launches,
            "avg_ms": avg, "algorithmic_bytes_per_launch": kp_bytes}
    cpu = None
    if args.mechanics_cpu_sample > 0:
        m = min(args.mechanics_cpu_sample, n)
        log(f"[bench] mechanics: the oracle on the {m}^3 system")
        cb = mechanics_cpu_sample(m, args.seed)
        scale = (cb["unknowns"] * cb["rounds"]) / (s.n * rounds) if m != n else 1.0
        cpu = {"value": scale / cb["seconds"], "unit": f"steps/sec ({n}^3" + (", scaled from the sample)" if m != n else ")"),
               "cores": 1, "kind": "port",
               "sample": f"oracle (C++ FP64 restatement of mechanics/src/solve.rs, one thread as the reference's "
                         f"Solver) Solver::new + solve to 1e-3 on the {m}^3 system of the same generator: "
                         f"{cb['seconds']:.2f} s, {cb['rounds']} rounds, {cb['unknowns']} unknowns"
                         + (f"; scaled by unknowns x rounds to the {n}^3 system's {s.n} x {rounds}" if m != n else "")}
    out = {
        "metric": f"mechanics solver steps/sec (Solver::new + solve to 1e-3) for the {n}^3 structure's system",
        "value": 1e3 / ms_step, "unit": "steps/sec", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"the bench generator's {n}^3 pillar/slab structure as co-rotated wood cubes of "
                               f"0.125 m (turned up to 0.2 rad), the held cells' corners held, dt 1/60: "
                               f"(M + (dt^2 + 0.3 dt) K) v' = M v + dt M g from random speeds",
                   "lattice": n, "cells": cells, "unknowns": int(s.n), "pairs": int(len(s.ij))},
        "rounds": rounds, "solve_ms": sum(solve_ms) / len(solve_ms), "setup_ms": info["setup_ms"],
        "levels": [lv[0] for lv in info["levels"]], "tight": {"tol": 1e-8, "rounds": rounds_tight, "ms": tight_ms},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--spacing", type=float, default=7.0,
                    help="pillar Poisson-disk spacing (7: ~29%% solid; 12.8: ~14%%, the 512^3 sparse world)")
    ap.add_argument("--points", type=int, default=64, help="point loads on top of gravity (1024: the 512^3 world)")
    ap.add_argument("--cpu-sample", type=int, default=64, help="edge of the CPU oracle's sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-full", type=int, default=0, metavar="N",
                    help="run the CPU oracle alone on the whole N^3 workload (progress lines, one JSON line)")
    ap.add_argument("--cpu-max-iter", type=int, default=100000, help="cpu-full: the oracle's iteration cap")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-errors", dest="errors", action="store_false",
                    help="skip the error report (displacements and gradient against tight FP64 solves)")
    ap.add_argument("--batch", type=int, default=0,
                    help="instead: the batch workload, this many structures (512: the batch config)")
    ap.add_argument("--batch-cells", type=float, default=24e6,
                    help="batch: free cells per st_solve_batch call (the structures are packed in order)")
    ap.add_argument("--fd-check", type=int, default=0, metavar="CELLS",
                    help="instead: the gradient check, d(compliance) and d(soft-max stress)/d(log-stiffness) "
                         "against central differences (eps 1e-4, FP64 solves to 1e-10) at this many cells "
                         "of a --size structure (64 at 128: the adjoint config)")
    ap.add_argument("--ingest", type=int, default=0, metavar="N",
                    help="instead: solids + hold_bulk from a lattice::Grid of one N^3 tile (st_grid_cells)")
    ap.add_argument("--gather", type=int, default=0, metavar="N",
                    help="instead: gather::gathered on the device for the N^3 structure (see gather_bench)")
    ap.add_argument("--gather-cpu-sample", type=int, default=96, help="gather: edge of the CPU oracle's sample")
    ap.add_argument("--settle", type=int, default=0, metavar="N",
                    help="instead: bodies::settle on the device for one N^3 structure (see settle_bench)")
    ap.add_argument("--settle-cpu-sample", type=int, default=32, help="settle: edge of the CPU oracle's sample")
    ap.add_argument("--settle-quantile", type=float, default=0.999,
                    help="settle: every strength at this quantile of the standing structure's used shares")
    ap.add_argument("--mechanics", type=int, default=0, metavar="N",
                    help="instead: mechanics::solve::Solver on the device for the N^3 structure's system")
    ap.add_argument("--mechanics-cpu-sample", type=int, default=48, help="mechanics: edge of the oracle's sample")
    ap.add_argument("--short-solve", type=int, default=0, metavar="ITERS",
                    help="upload the structure, time every operator schedule and run a mixed solve of ITERS "
                         "iterations (for counter runs: every operator kernel and epilogue, briefly)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="end with three device copies of 1 GiB (the calibration kernel of the traffic counters)")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="offline: HBM bytes per launch of the roofline kernel from two rocprofv3 --pmc "
                         "counter_collection.csv files (FETCH_SIZE, WRITE_SIZE) of a --copy-calib run; "
                         "writes " + traffic_file("<kernel>"))
    ap.add_argument("--traffic-schedule", type=int, default=None,
                    help="with --traffic: the STATICS_HIP_SCHEDULE the counter runs were pinned to")
    args = ap.parse_args()
    if args.traffic:
        traffic_summary(*args.traffic, schedule=args.traffic_schedule)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path is single-GPU by design", file=sys.stderr)
        sys.exit(2)
    if args.cpu_full:
        cpu_full(args)
        return
    if args.short_solve:
        short_solve(args)
        return
    if args.batch:
        batch_bench(args)
        return
    if args.fd_check:
        fd_check(args)
        return
    if args.settle:
        settle_bench(args)
        return
    if args.mechanics:
        mechanics_bench(args)
        return
    if args.gather:
        gather_bench(args)
        return
    if args.ingest:
        ingest_bench(args)
        return
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731

    n = args.size
    t = time.perf_counter()
    occ, theta = structure(n, args.seed, spacing=args.spacing)
    sites, held, e, g = sh.cells_from_occ(occ, theta)
    log(f"[bench] generated {n}^3: {len(sites)} solid cells, fill {len(sites) / n ** 3:.3f} "
        f"({time.perf_counter() - t:.1f}s)")
    t = time.perf_counter()
    ctx = sh.Context().upload(sites, held, e, g)
    log(f"[bench] uploaded: {ctx.nodes} nodes, {ctx.cells} free cells, {ctx.bricks} bricks "
        f"({time.perf_counter() - t:.1f}s)")
    joined = ctx.joined_to_held()
    faces = ctx.node_faces()
    nheld = ctx.node_held()
    loads = node_loads(occ, faces, joined, nheld, np.random.default_rng(args.seed), n_points=args.points)
    loads = np.ascontiguousarray(loads)
    ctx._check(ctx.L.st_set_loads(ctx.h, sh._p(loads)), "st_set_loads")

    import ctypes
    # the north star's mode: each solve frees its hierarchy, the adjoint recomputes it
    opts_mixed = sh.StSolveOpts(sh.MIXED, 1e-8, 20000, 0, 0, 1)
    opts_fp64 = sh.StSolveOpts(sh.FP64, 1e-8, 20000, 0, 0, 1)

    def classify():
        c = ctypes.c_int64()
        ctx._check(ctx.L.st_classify(ctx.h, ctypes.byref(c)), "st_classify")
        return c.value

    def solve(opts, check=True):
        s = sh.StStats()
        rc = ctx.L.st_solve_resident(ctx.h, ctypes.byref(opts), ctypes.byref(s))
        if check or rc != sh.ST_NO_CONVERGENCE:
            ctx._check(rc, "st_solve_resident")
        return s

    def adjoint(opts, dJ=None, check=True):
        s = sh.StStats()
        J = ctypes.c_double()
        rc = ctx.L.st_backward(ctx.h, sh.SOFTMAX_STRESS, 20.0, ctypes.byref(opts), ctypes.byref(J),
                               None if dJ is None else sh._p(dJ), ctypes.byref(s))
        if check or rc != sh.ST_NO_CONVERGENCE:
            ctx._check(rc, "st_backward")
        return J.value, s

    def step(opts):
        classify()
        a = solve(opts)
        J, b = adjoint(opts)
        return a, b, J

    def solution():
        x = np.zeros((ctx.nodes, 6))
        ctx._check(ctx.L.st_get_solution(ctx.h, sh._p(x)), "st_get_solution")
        return x

    for i in range(args.warmup):
        a, b, J = step(opts_mixed)
        log(f"[bench] warmup {i}: forward {a.iters} it ({a.setup_ms:.1f}+{a.solve_ms:.1f} ms, res {a.rel_residual:.2e}), "
            f"adjoint {b.iters} it ({b.setup_ms:.1f}+{b.solve_ms:.1f} ms, res {b.rel_residual:.2e})")
    ctx._check(ctx.L.st_kernel_timing_enable(ctx.h, 0, 1), "timing")
    times, fwd, adj = [], [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        a, b, J = step(opts_mixed)
        times.append(time.perf_counter() - t0)
        fwd.append(a)
        adj.append(b)
        log(f"[bench] step {i}: {times[-1] * 1e3:.1f} ms; forward {a.iters} it ({a.setup_ms:.0f}+{a.solve_ms:.0f} ms), "
            f"res {a.rel_residual:.2e}; adjoint {b.iters} it ({b.setup_ms:.0f}+{b.solve_ms:.0f} ms), "
            f"res {b.rel_residual:.2e}; J {J:.6g}")
    nl, tms, bpl = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    ctx._check(ctx.L.st_kernel_timing(ctx.h, ctypes.byref(nl), ctypes.byref(tms), ctypes.byref(bpl)), "timing")
    ctx._check(ctx.L.st_kernel_timing_enable(ctx.h, 0, 0), "timing")
    kidx, kname, kdesc, kdisp = ctx.apply_schedule(fp64=True)
    kernel_ms = tms.value / max(nl.value, 1)
    achieved = bpl.value / (kernel_ms * 1e-3) / 1e9 if nl.value else 0.0
    ms_per_step = 1e3 * sum(times) / len(times)
    dof = 6 * fwd[-1].dof_nodes

    # the whole PCG iteration against the HBM roofline: the canonical 25 vector passes of
    # the reference's algorithm per iteration on the fine level, 11 of them FP64 (the CG)
    # and 14 FP32 (the V-cycle) in the mixed mode: 144 D bytes; 200 D in plain FP64
    def iteration_roofline(stats_list, bytes_per_dof):
        its = sum(s.iters for s in stats_list)
        ms = sum(s.solve_ms for s in stats_list)
        per_it = ms / max(its, 1)
        gbs = bytes_per_dof * dof / (per_it * 1e-3) / 1e9
        return {"bound": "hbm", "ms_per_iteration": per_it, "bytes_per_iteration": bytes_per_dof * dof,
                "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
                "count": f"{bytes_per_dof} B per DOF per iteration (25 canonical vector passes), D = {dof}; "
                         "coarse levels excluded; iteration time = solve time / iterations, forward and adjoint"}

    rl_iter_mixed = iteration_roofline([fwd[-1], adj[-1]], 144)

    # the same step in plain FP64, like for like (recomputed hierarchy in both solves)
    fp64_step = None
    if not args.no_fp64:
        step(opts_fp64)  # warm
        t0 = time.perf_counter()
        a64, b64, J64 = step(opts_fp64)
        fp64_step = dict(ms=(time.perf_counter() - t0) * 1e3, fwd_iters=a64.iters, adj_iters=b64.iters,
                         fwd_setup_ms=a64.setup_ms, adj_setup_ms=b64.setup_ms, J=J64,
                         iteration_roofline=iteration_roofline([a64, b64], 200))
        log(f"[bench] fp64 step: {fp64_step['ms']:.1f} ms; forward {a64.iters} it, adjoint {b64.iters} it")

    # the forward solves alone (classify + solve), and their errors against a tight FP64 solve
    t0 = time.perf_counter()
    classify()
    m1 = solve(opts_mixed)
    mixed_fwd_ms = (time.perf_counter() - t0) * 1e3
    x_mixed = solution()
    fp64 = None
    if not args.no_fp64:
        t0 = time.perf_counter()
        classify()
        p = solve(opts_fp64)
        fp64 = dict(ms=(time.perf_counter() - t0) * 1e3, iters=p.iters, rel_residual=p.rel_residual)
        x_fp64 = solution()
        log(f"[bench] fp64 forward: {fp64['ms']:.1f} ms, {p.iters} it, res {p.rel_residual:.2e}")
    log(f"[bench] mixed forward: {mixed_fwd_ms:.1f} ms, {m1.iters} it, res {m1.rel_residual:.2e}")
    errors = None
    if args.errors:
        # reference: FP64 as far as its true residual goes in 400 iterations (it floors near 4e-10)
        log("[bench] errors: FP64 reference solve (400 iterations at most)")
        ref = solve(sh.StSolveOpts(sh.FP64, 1e-12, 400, 0, 0, 0), check=False)
        log(f"[bench] errors: reference {ref.iters} it, res {ref.rel_residual:.2e}")
        x_ref = solution()
        scale = np.abs(x_ref).max()
        errors = {"reference": f"FP64 device solve, {ref.iters} iterations, true residual {ref.rel_residual:.2e}",
                  "mixed_forward_max_rel": float(np.abs(x_mixed - x_ref).max() / scale)}
        if fp64:
            errors["fp64_forward_max_rel"] = float(np.abs(x_fp64 - x_ref).max() / scale)
        del x_ref
        # the soft-max gradient of the mixed step against FP64 solves to 1e-10 (or as far as
        # their true residuals go: at 512^3 they floor above it)
        g_ref = np.zeros(len(sites))
        rf = solve(sh.StSolveOpts(sh.FP64, 1e-10, 600, 0, 0, 0), check=False)
        log(f"[bench] errors: gradient reference forward {rf.iters} it, res {rf.rel_residual:.2e}")
        _, ra = adjoint(sh.StSolveOpts(sh.FP64, 1e-10, 600, 0, 0, 0), g_ref, check=False)
        log(f"[bench] errors: gradient reference adjoint {ra.iters} it, res {ra.rel_residual:.2e}")
        g_mix = np.zeros(len(sites))
        solve(opts_mixed)
        adjoint(opts_mixed, g_mix)
        errors["mixed_softmax_gradient_max_rel"] = float(np.abs(g_mix - g_ref).max() / np.abs(g_ref).max())
        errors["gradient_reference"] = (f"FP64 forward and adjoint solves to 1e-10 (true residuals reached "
                                        f"{rf.rel_residual:.2e} and {ra.rel_residual:.2e})")
        log(f"[bench] errors: {errors}")
    del x_mixed

    cpu = None
    if args.cpu_sample > 0:
        log(f"[bench] cpu oracle on a {args.cpu_sample}^3 sample, {args.cpu_threads} threads")
        cb = cpu_baseline(args.cpu_sample, args.cpu_threads, args.seed, args.points, spacing=args.spacing)
        what = (f"oracle (C++ FP64 restatement of statics/src/bodies.rs: assembled blocks, smoothed aggregation, "
                f"Cholesky top) classify + motions to 1e-8 on the {args.cpu_sample}^3 structure of the same generator "
                f"and loads: {cb['seconds']:.2f} s ({cb['setup_s']:.2f} s setup), {cb['iters']} iterations, "
                f"{cb['dof']} DOF")
        full = cpu_full_record(n, args)
        ref = cpu_full_record(args.cpu_sample, args)
        if full and ref and not full["capped"]:
            # the whole workload's time was measured (the oracle on every DOF to 1e-8, on this kind of
            # host): the live sample is scaled by the measured ratio of the two runs
            factor = full["seconds"] / ref["seconds"]
            per_s = 1.0 / (cb["seconds"] * factor)
            unit = f"structure solves/sec ({n}^3; the live sample times the measured {n}^3 / sample ratio)"
            what += (f"; the oracle on the whole {n}^3 workload ({full['dof']} DOF, {full['iters']} iterations, "
                     f"{full['seconds']:.0f} s: {full['setup_s']:.0f} s setup, {full['solve_s']:.0f} s solve, "
                     f"{full['maxrss_gb']:.0f} GB, {full['threads']} threads, {full['cpu']}) against "
                     f"{ref['seconds']:.2f} s for the sample in the same run set: ratio {factor:.1f} "
                     f"(profiles/cpu_oracle_{n}.txt)")
        else:
            per_s = (1.0 / cb["seconds"]) * (cb["dof"] / dof)
            unit = f"structure solves/sec ({n}^3, scaled linearly by DOF from the sample)"
            what += (f"; scaled linearly by DOF to the {n}^3 structure's {dof} DOF, which flatters the CPU: its "
                     f"iteration count grows with the size")
        cpu = {"value": per_s, "unit": unit, "cores": args.cpu_threads, "kind": "port", "cpu_model": cb["cpu"],
               "nproc": os.cpu_count(), "sample": what}
        log(f"[bench] cpu sample: {cb}")

    value = 1e3 / ms_per_step
    out = {
        "metric": f"structure solves/sec to 1e-8 residual at {n}^3; mixed+adjoint vs plain FP64",
        "value": value,
        "unit": "structure solves/sec",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 CG + f32 V-cycle",
        "data": "synthetic",
        "config": {"workload": f"one {n}^3 pillar/slab structure, {len(sites) / n ** 3:.0%} solid (pillars 2-6 wide "
                               f"at Poisson spacing {args.spacing:g}), lognormal stiffness sigma 0.5, gravity + "
                               f"{args.points} point loads; step = classify + mixed solve + soft-max stress adjoint, "
                               f"each solve recomputing its multigrid hierarchy",
                   "lattice": n, "free_cells": int(fwd[-1].cells), "free_nodes": int(fwd[-1].dof_nodes),
                   "dof": int(dof), "tol": 1e-8, "hierarchy": "recomputed per solve (release_hierarchy)"},
        "forward_iters": int(fwd[-1].iters), "adjoint_iters": int(adj[-1].iters),
        "forward_setup_ms": fwd[-1].setup_ms, "forward_solve_ms": fwd[-1].solve_ms,
        "adjoint_setup_ms": adj[-1].setup_ms, "adjoint_solve_ms": adj[-1].solve_ms,
        "forward_rel_residual": fwd[-1].rel_residual, "adjoint_rel_residual": adj[-1].rel_residual,
        "device_memory_peak_gb": max(fwd[-1].mem_peak, adj[-1].mem_peak) / 1e9,
        "fp64_step_ms": fp64_step["ms"] if fp64_step else None,
        "mixed_vs_fp64_step": (fp64_step["ms"] / ms_per_step) if fp64_step else None,
        "fp64_step": fp64_step,
        "mixed_forward_solves_per_s": 1e3 / mixed_fwd_ms,
        "fp64_forward_solves_per_s": (1e3 / fp64["ms"]) if fp64 else None,
        "fp64_forward_iters": fp64["iters"] if fp64 else None,
        "mixed_vs_fp64_forward": (fp64["ms"] / mixed_fwd_ms) if fp64 else None,
        "errors": errors,
        "roofline": {"bound": "hbm",
                     "kernel": f"{kname}<double, EpiDot<double>> (the CG's K p with the fused p.Ap, FP64; "
                               f"{kdesc}; {kdisp} dispatch(es) per launch)", "schedule": kidx,
                     "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     # the counter runs were of the default 256^3 workload
                     "traffic": measured_traffic(kname) if (n, args.spacing, args.seed) == (256, 7.0, 1) else None,
                     "launches": int(nl.value), "avg_ms": kernel_ms,
                     "dispatches_per_launch": kdisp, "algorithmic_bytes_per_launch": bpl.value},
        "roofline_iteration": rl_iter_mixed,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if args.copy_calib:
        copy_calibration()


if __name__ == "__main__":
    main()
