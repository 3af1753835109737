#!/usr/bin/env python3
"""Benchmark of the D3Q27 entropic collide-stream step on one MI355X.

Default workload (the configuration the metric is quoted on): a 256^3 duct,
inlet at Ma 0.2 on x-, equilibrium pressure outlet (rho = 1) on x+,
bounce-back walls on y and z, and a level-5 icosphere (20,480 triangles,
D = 48, centre (80.5, 128.5, 128.5), rotation seed 20260102) as the solid,
its wall links rebuilt by interpolated bounce-back.  Link tables are built on
the device before timing.  One step = the fused pull-stream + moments +
entropic alpha (FP64 Newton) + relaxation kernel over the interior nodes plus
the boundary-node kernel, populations FP32 stored as deviations from the
lattice weights, already resident in HBM.  `--workload tgv` runs the periodic
Taylor-Green box instead, `--workload c3` the D3Q39 two-population
compressible step of the supersonic wedge case (384 x 192 x 96, Mach 1.5,
shifted lattice, 10 degree wedge with the shifted-lattice wall rule, 2000
steps with the shock angle measured at the end), reported against its
algorithmic 728 B per node update.

Prints ONE JSON line.  `value` is MLUPS over the timed
region (wall clock, synchronised on both sides); `roofline` uses the
algorithmic 216 B per node update (27 populations x 4 B, read + write) and the
kernel's average duration from HIP events recorded on the library's stream.
`cpu_baseline` times the FP64 oracle on this host's cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "university-of-cincinnati-lattice-boltzmann-solver_amd")
HBM_PEAK_GBS = 8000.0
BYTES_PER_NODE = 27 * 4 * 2


def load_pkg():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "uclbs_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["uclbs_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def tgv_params(L, mach=0.1, re=1600.0):
    u0 = mach / math.sqrt(3.0)
    return u0, u0 * (L / (2 * math.pi)) / re


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(budget_s: float):
    """FP64 oracle (C/OpenMP restatement of the path) on a 64^3 TGV sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # test infrastructure: the CPU baseline leg only
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    threads = max(1, min(threads, 16))
    L = 64
    shape = (L, L, L)
    u0, nu = tgv_params(L)
    beta = 1 / (6 * nu + 1)
    f = oracle.init_tgv27(shape, u0, 1e-4, 20260101)
    # one step to size the sample, then as many as fit the budget
    t0 = time.perf_counter()
    f, _, _ = oracle.step_periodic27(shape, f, beta, nsteps=1, nthreads=threads, want_alpha=False)
    dt1 = time.perf_counter() - t0
    steps = max(2, min(200, int(budget_s / max(dt1, 1e-6))))
    t0 = time.perf_counter()
    oracle.step_periodic27(shape, f, beta, nsteps=steps, nthreads=threads, want_alpha=False)
    dt = time.perf_counter() - t0
    return {"value": round(L ** 3 * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": threads,
            "kind": "port",
            "sample": f"TGV {L}^3 D3Q27 FP64 entropic Newton, {steps} steps, "
                      f"{threads} OpenMP threads on {cpu_model()}"}


def cpu_baseline_c2(s, L, nu, budget_s: float):
    """FP64 oracle (C/OpenMP restatement of the path) on the same C2 duct: the
    device's link table and its current populations, stepped by the oracle's
    box step (faces, inlet/outlet equilibria, wall links, entropic Newton) for
    as many steps as fit the budget (at least 2)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # test infrastructure: the CPU baseline leg only
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        threads = os.cpu_count() or 1
    threads = max(1, min(threads, 16))
    u_in = 0.2 / math.sqrt(3.0)
    state = s.download_links(0)
    f = s.download_f(0)
    kw = dict(inlets=[(1.0, u_in, 0.0, 0.0)], outlets=[(1.0, u_in, 0.0, 0.0)], tol=1e-10,
              nthreads=threads, want_alpha=False)
    beta = 1 / (6 * nu + 1)
    shape = (L, L, L)
    t0 = time.perf_counter()
    f, _, _, _ = oracle.step_box27(shape, C2_FACES, state, f, beta, nsteps=1, **kw)
    dt1 = time.perf_counter() - t0
    steps = max(2, min(50, int(budget_s / max(dt1, 1e-6))))
    t0 = time.perf_counter()
    oracle.step_box27(shape, C2_FACES, state, f, beta, nsteps=steps, **kw)
    dt = time.perf_counter() - t0
    return {"value": round(L ** 3 * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": threads,
            "kind": "port",
            "sample": f"C2 {L}^3 duct with the sphere (the device's link table and state after "
                      f"the timed steps), D3Q27 FP64 entropic Newton, {steps} steps, "
                      f"{threads} OpenMP threads on {cpu_model()}"}


FP64_PEAK_TFS = 78.6  # MI355X FP64 vector / matrix peak
# FP64 flops per D3Q39 node update = base + per Newton iteration, from
# SQ_INSTS_VALU_FLOPS_FP64 over the three k39_step launches of a C3 step at two
# points of the run (1.10 iterations per node: 13.6k; 2.5: 27.0k)
C3_FLOPS_BASE, C3_FLOPS_PER_ITER = 3.16e3, 9.52e3


def read_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    try:
        with open(p) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return None


C2_FACES = [-101, -10, -10, -201, -10, -10]  # inlet x-, walls y/z, outlet x+


def make_c2(pkg, L, nu, alpha_mode):
    """The C2 duct: faces, inlet/outlet states, icosphere link tables, uniform start."""
    u_in = 0.2 / math.sqrt(3.0)
    s = pkg.Solver(precision="fp32s", nu=nu, alpha_mode=alpha_mode, alpha_tol=1e-10)
    s.define_level(0, (L, L, L), periodic=(0, 0, 0))
    s.set_faces(0, C2_FACES)
    s.set_bc_states(inlets=[(1.0, u_in, 0.0, 0.0)], outlets=[(1.0, u_in, 0.0, 0.0)])
    sc = L / 256.0
    tri, nrm = pkg.icosphere(5, 24.0 * sc, (80.5 * sc, 128.5 * sc, 128.5 * sc), 20260102)
    n_solid, n_tie = s.build_links(0, tri, nrm)
    n_bnd, _ = s.link_counts(0)
    s.init_uniform(0, 1.0, (u_in, 0.0, 0.0))
    return s, n_solid, n_tie, n_bnd


def learned_alpha(pkg, s, L, nu, collect_steps=3, train_steps=4000, batch=8192, steps=20):
    """Second part of the metric: Newton updates per node of the plain start
    (alpha = 2) and of the learned start at the same state.  The network is
    trained on pairs logged from this plain run (every interior node with
    max|x| > 1e-8 of `collect_steps` successive steps), then both forms advance
    the same state `steps` steps.  Not timed in `value`."""
    import numpy as np
    t0 = time.perf_counter()
    s.mlp_reserve(collect_steps * L ** 3)
    npairs = 0
    for _ in range(collect_steps):
        npairs = s.mlp_collect(0, 1)
    t_collect = time.perf_counter() - t0
    s.mlp_init(20260105)
    t0 = time.perf_counter()
    l0, l1 = s.mlp_train(train_steps, batch, 2e-3, 20260105)
    t_train = time.perf_counter() - t0
    params = s.mlp_get_params()
    f_state = s.download_f(0)
    sl, _, _, _ = make_c2(pkg, L, nu, "learned")
    sl.mlp_set_params(*params)
    sl.upload_f(0, f_state)
    del f_state
    out = {}
    for name, solver in (("plain", s), ("learned", sl)):
        solver.reset_stats()
        solver.sync()
        t0 = time.perf_counter()
        solver.step(0, steps)
        dt = time.perf_counter() - t0
        st = solver.stats()
        out[name] = {"newton_iters_per_node": st["newton_iters"] / st["node_updates"],
                     "fallback_rate": st["n_fallback"] / st["node_updates"],
                     "mlups": L ** 3 * steps / dt / 1e6}
    mp, ml = s.macro(0), sl.macro(0)
    up, ul = np.stack(mp[1:]), np.stack(ml[1:])
    rel = float(np.linalg.norm(ul - up) / np.linalg.norm(up))
    ip, il = out["plain"]["newton_iters_per_node"], out["learned"]["newton_iters_per_node"]
    finite = all(math.isfinite(v) for v in (l0, l1, ip, il, rel))
    return {"status": "ok" if finite else "failed",
            "plain_iters_per_node": round(ip, 4), "learned_iters_per_node": round(il, 4),
            "iters_ratio": round(ip / il, 3), "learned_fallback_rate": out["learned"]["fallback_rate"],
            "u_rel_l2_learned_vs_plain": rel, "steps": steps, "train_pairs": npairs,
            "train_steps": train_steps, "batch": batch, "loss_first": l0, "loss_last": l1,
            "collect_s": round(t_collect, 2), "train_s": round(t_train, 2),
            "plain_mlups": round(out["plain"]["mlups"], 1),
            "learned_mlups": round(out["learned"]["mlups"], 1)}


C3_SHAPE = (384, 192, 96)
C3_FACES = [-1000, -1000, -11, -6, -1000, -11]   # freestream x-/y-/y+, outflow x+, inviscid z
C3_APEX, C3_LENGTH, C3_HALF_ANGLE, C3_SEED = (96.3, 96.27), 64.0, 10.0, 20260103
C3_STEPS = 2000


def bench_c3(pkg, args):
    """D3Q39 supersonic wedge: 384 x 192 x 96 shifted lattice at Mach 1.5, a
    10 degree half-angle wedge (seeded triangulation, 37,248 triangles, apex
    (96.3, 96.27), length 64, spanning z) with its wall links rebuilt by the
    shifted-lattice wall rule every step, freestream inflow, outflow, inviscid
    sidewalls; FP32 deviations from the freestream equilibrium, FP64
    multipliers.  From the uniform start: the 2000-step run with its last
    `steps` steps timed (at least `warmup` before them); then the shock angle
    and shock-layer density on the z mid-plane against theta-beta-Mach."""
    import importlib
    sh = importlib.import_module("uclbs_amd.shock")
    shape = C3_SHAPE
    n = shape[0] * shape[1] * shape[2]
    T0 = 1 / 1.4
    t_setup = time.perf_counter()
    s = pkg.Solver(lattice="d3q39", precision="fp32s")
    s.set_gas(mach=1.5, shifted=1, turbulent=True)
    s.define_level(0, shape, periodic=(0, 0, 0))
    s.set_faces(0, C3_FACES)
    tri, nrm = pkg.wedge(C3_APEX, C3_LENGTH, C3_HALF_ANGLE, (-5.0, shape[2] + 5.0), 64, 96,
                         C3_SEED)
    n_solid, _ = s.build_links(0, tri, nrm)
    s.init_fields(0, 1.0, 1.5, 0.0, 0.0, T0)
    s.set_timestep(0, 0)
    setup_s = time.perf_counter() - t_setup
    # the timed steps are the last ones of the run, on the developed flow (the
    # start-up near the uniform state needs half the Newton iterations)
    before = max(args.warmup, args.c3_total - args.steps)
    s.step(0, before)
    s.reset_stats()
    s.sync()
    t0 = time.perf_counter()
    step_ms, main_ms = s.time_kernels(0, args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    st = s.stats()
    rho = s.macro39(0)[0].reshape(shape[2], shape[1], shape[0])[shape[2] // 2]
    solid = (s.download_links(0)[0] < 0).reshape(shape[2], shape[1], shape[0])[shape[2] // 2]
    beta, rho_rh, _ = sh.oblique_shock(1.5, C3_HALF_ANGLE)
    ang, post, ncol = sh.shock_fit(rho, solid, C3_APEX, C3_HALF_ANGLE, (8, 56))
    bpn = 2 * 39 * 4 * 2 + 13 * 4 * 2
    avg = main_ms / 1e3 / args.steps
    fpn = C3_FLOPS_BASE + C3_FLOPS_PER_ITER * st["newton_iters"] / max(1, st["node_updates"])
    achieved = bpn * n / avg / 1e9
    out = {"metric": "MLUPS, D3Q39 supersonic wedge step (secondary workload)",
           "value": round(n * args.steps / wall / 1e6, 1), "unit": "MLUPS", "n_gpus": 1,
           "steps": args.steps, "warmup": before,
           "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 populations (shifted), f64 multipliers/moments/alpha",
           "data": "synthetic",
           "config": {"workload": "D3Q39 384x192x96, Mach 1.5 shifted lattice, 10 degree wedge "
                                  "(seed 20260103), freestream inflow x- and far field y, "
                                  "outflow x+, inviscid z sidewalls, uniform start",
                      "nodes": n, "solid_nodes": n_solid, "lattice": "D3Q39",
                      "setup_s": round(setup_s, 2), "total_steps": before + args.steps},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": None,
                        "kernel": "one D3Q39 step: wall rebuild, k39_step (interior, near-face "
                                  "and wall nodes), repair pass",
                        "avg_launch_ms": round(avg * 1e3, 4),
                        # the step is bound by FP64 VALU work: FP64 flops per node update
                        # from the timed steps' Newton iterations (counted model below)
                        "fp64": {"flops_per_node": round(fpn),
                                 "achieved": round(fpn * n / avg / 1e12, 2),
                                 "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                                 "frac": round(fpn * n / avg / 1e12 / FP64_PEAK_TFS, 4)}},
           "shock": {"angle_deg": round(ang, 3), "theory_deg": round(beta, 3),
                     "layer_rho": round(post, 4), "rankine_hugoniot_rho": round(rho_rh, 4),
                     "columns": ncol, "at_step": before + args.steps},
           "lm_stats": {"newton_iters_per_node": round(st["newton_iters"] / max(1, st["node_updates"]), 4),
                        "n_noconv": st["n_noconv"], "n_repair": st["n_repair"],
                        "n_wall_links": st["n_ibb_links"],
                        "n_alpha_trivial": st["n_trivial"], "n_alpha_bound": st["n_bound"]}}
    print(json.dumps(out), flush=True)


def _split(lo, hi, edges, rng):
    out, a = [], lo
    while a <= hi:
        b = min(hi, a + int(rng.choice(edges)) - 1)
        if hi - b < min(edges):
            b = hi
        out.append((a, b))
        a = b + 1
    return out


def _tiling(lo, hi, edges, rng):
    """Uneven boxes tiling lo..hi (inclusive): x segments drawn anew per (z, y) slab."""
    boxes = []
    for z0, z1 in _split(lo[2], hi[2], edges, rng):
        for y0, y1 in _split(lo[1], hi[1], edges, rng):
            for x0, x1 in _split(lo[0], hi[0], edges, rng):
                boxes.append(((x0, y0, z0), (x1, y1, z1)))
    return boxes


def c4_levels(L=256, seed=20260104):
    """Levels 1 and 2 of the C2 geometry as uneven boxes (ratio 2): level 1
    covers the sphere and its near wake (coarse x 40-152, y and z 88-168;
    edges 16-64), level 2 the sphere's surroundings (coarse x 52-110, y and z
    100-157; edges 8-48), both seeded; level 0 is the C2 duct itself."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sc = L / 256.0
    c = lambda v: int(round(v * sc))  # noqa: E731
    l1 = _tiling((2 * c(40), 2 * c(88), 2 * c(88)), (2 * c(152) - 1, 2 * c(168) - 1, 2 * c(168) - 1),
                 (16, 24, 32, 48, 64), rng)
    l2 = _tiling((4 * c(52), 4 * c(100), 4 * c(100)), (4 * c(110) - 1, 4 * c(157) - 1, 4 * c(157) - 1),
                 (8, 16, 24, 32, 48), rng)
    return [l1, l2]


def bench_c4(pkg, args):
    """Batched AMR on the C2 geometry: the C2 duct (one box) with two refined
    levels of uneven boxes around the sphere, each carrying its device-built
    link table; all boxes of a level in one launch, the two-substep level
    schedule in the library replayed as a hipGraph; MLUPS over all levels."""
    import numpy as np
    L = args.size
    nu = args.nu  # the C2 viscosity on level 0 (doubled per level, acoustic scaling)
    t_setup = time.perf_counter()
    s, n_solid, _, _ = make_c2(pkg, L, nu, "newton")
    sc = L / 256.0
    tri, nrm = pkg.icosphere(5, 24.0 * sc, (80.5 * sc, 128.5 * sc, 128.5 * sc), 20260102)
    levels = c4_levels(L)
    for lvl, boxes in enumerate(levels, start=1):
        s.define_level_boxes(lvl, boxes)
        s.build_links(lvl, tri, nrm, origin=(0.0, 0.0, 0.0), dx=0.5 ** lvl)
        s.init_from_coarse(lvl)
    tiles = [s.autotune(lvl) for lvl in range(1, 3)]
    setup_s = time.perf_counter() - t_setup
    valid = [L ** 3] + [sum(int(np.prod([h - l + 1 for l, h in zip(*bx)])) for bx in boxes)
                        for boxes in levels]
    upd = sum(v << lvl for lvl, v in enumerate(valid))  # level l: 2^l substeps per coarse step
    warm = max(2, args.warmup // 10)
    K = max(2, args.steps // 10) // 2 * 2
    s.advance_native(warm, graph=True)
    s.reset_stats()
    s.sync()
    t0 = time.perf_counter()
    s.advance_native(K, graph=True)
    wall = time.perf_counter() - t0
    st = s.stats()
    t1 = time.perf_counter()
    s.advance_native(K, graph=False)
    wall_eager = time.perf_counter() - t1
    # the dominant kernel: level 2's box kernel (4 substeps per coarse step)
    step_ms, main_ms = s.time_kernels(2, 4)
    avg2 = main_ms / 1e3 / 4
    ach2 = 216.0 * valid[2] / avg2 / 1e9
    rho = s.macro(0)[0]
    out = {"metric": "MLUPS over all levels, D3Q27 3-level batched AMR (secondary workload)",
           "value": round(upd * K / wall / 1e6, 1), "unit": "MLUPS", "n_gpus": 1,
           "steps": K, "warmup": warm, "ms_per_step": round(wall * 1e3 / K, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 populations (shifted), f64 moments/alpha", "data": "synthetic",
           "config": {"workload": f"C4: the C2 duct {L}^3 (inlet, outlet, walls, sphere of "
                                  f"20,480 triangles) with 2 refined levels of uneven boxes "
                                  f"around the sphere ({', '.join(str(len(b)) for b in levels)} "
                                  f"boxes, edges 8-64, seeded), link tables on every level, "
                                  f"entropic alpha (newton), FP32-shifted; one batched launch per "
                                  f"level step, two coarse steps per hipGraph replay",
                      "valid_nodes_per_level": valid, "updates_per_coarse_step": upd,
                      "solid_nodes_level0": n_solid, "tile_nodes": tiles,
                      "setup_s": round(setup_s, 2),
                      "eager_mlups": round(upd * K / wall_eager / 1e6, 1),
                      "rho_range_level0": [round(float(np.nanmin(rho)), 5), round(float(np.nanmax(rho)), 5)]},
           "roofline": {"bound": "hbm", "achieved": round(ach2, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(ach2 / HBM_PEAK_GBS, 4), "traffic": None,
                        "kernel": "k27_step_boxes (level 2, all boxes in one launch)",
                        "avg_launch_ms": round(avg2 * 1e3, 4)},
           "alpha_stats": {"newton_iters_per_node": round(st["newton_iters"] / max(1, st["node_updates"]), 4),
                           "n_logpath": st["n_logpath"], "n_noconv": st["n_noconv"],
                           "n_ibb_links": st["n_ibb_links"]}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--workload", default="c2", choices=["c2", "tgv", "c3", "c4"])
    ap.add_argument("--c3-total", type=int, default=C3_STEPS,
                    help="total steps of the c3 run from the uniform start (the shock is "
                         "measured at the end; shorter runs serve profiling)")
    ap.add_argument("--nu", type=float, default=0.002, help="C2 viscosity (lattice units)")
    ap.add_argument("--alpha", default="newton", choices=["newton", "fixed2", "learned"],
                    help="C2: the timed form; the learned-vs-plain comparison runs after "
                         "the timing unless --no-learned")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the plain copy of one population buffer (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        nb = 27 * 4 * args.size ** 3
        pkg = load_pkg()
        print(json.dumps({"copy_bytes_each_way": nb,
                          "copy_GBps": round(pkg.copy_bandwidth(nb, reps=5), 1)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)

    pkg = load_pkg()
    if args.workload == "c3":
        return bench_c3(pkg, args)
    if args.workload == "c4":
        return bench_c4(pkg, args)
    L = args.size
    n = L ** 3
    # the timed form: plain entropic Newton (or alpha = 2); the learned start is
    # measured against it after the timing
    mode = "fixed2" if args.alpha == "fixed2" else "newton"
    t_setup = time.perf_counter()
    if args.workload == "c2":
        u_in = 0.2 / math.sqrt(3.0)
        nu = args.nu
        s, n_solid, n_tie, n_bnd = make_c2(pkg, L, nu, mode)
        workload = (f"C2: D3Q27 duct {L}^3, inlet Ma 0.2 (u = {u_in:.5f}), equilibrium outlet "
                    f"rho = 1, bounce-back walls y/z, level-5 icosphere (20480 triangles, "
                    f"D = {48 * L / 256.0:g}) with interpolated bounce-back, nu = {nu}, entropic alpha "
                    f"({mode}), FP32-shifted populations")
        extra = {"solid_nodes": n_solid, "boundary_nodes": n_bnd, "ties": n_tie}
    else:
        u0, nu = tgv_params(L)
        s = pkg.Solver(precision="fp32s", nu=nu, alpha_mode=mode, alpha_tol=1e-10)
        s.define_level(0, (L, L, L))
        s.init_tgv(0, u0, 1e-4, 20260101)
        workload = (f"periodic Taylor-Green D3Q27 {L}^3, Ma 0.1, Re 1600, entropic alpha "
                    f"({mode}), FP32-shifted populations")
        extra = {}
    setup_s = time.perf_counter() - t_setup
    s.step(0, args.warmup)
    s.reset_stats()
    s.sync()
    t0 = time.perf_counter()
    step_ms, main_ms = s.time_kernels(0, args.steps)  # HIP events on the library stream
    s.sync()
    wall = time.perf_counter() - t0
    st = s.stats()
    mlups = n * args.steps / wall / 1e6
    avg_launch_s = main_ms / 1e3 / args.steps
    achieved = BYTES_PER_NODE * n / avg_launch_s / 1e9
    achieved_step = BYTES_PER_NODE * n / (step_ms / 1e3 / args.steps) / 1e9
    traffic = read_traffic(f"traffic_d3q27_{args.workload}{L}.json")
    ceiling = pkg.copy_bandwidth(27 * 4 * n)  # plain copy of one population buffer
    out = {
        "metric": "MLUPS & %HBM roofline, D3Q27 256^3; learned vs plain alpha iters at equal err",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 populations (shifted), f64 moments/alpha",
        "data": "synthetic",
        "config": dict({"workload": workload, "nodes": n, "lattice": "D3Q27",
                        "alpha_mode": mode, "setup_s": round(setup_s, 2)}, **extra),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "kernel": (traffic or {}).get("kernel", "k27_step_periodic") +
                               " (streaming kernel, FP32 shifted)",
                     "achieved_whole_step": round(achieved_step, 1),
                     "frac_whole_step": round(achieved_step / HBM_PEAK_GBS, 4),
                     "measured_copy_GBps": round(ceiling, 1),
                     "frac_of_copy": round(achieved / ceiling, 4),
                     "avg_launch_ms": round(avg_launch_s * 1e3, 4)},
        "alpha_stats": {"newton_iters_per_node": round(st["newton_iters"] / max(1, st["node_updates"]), 4),
                        "n_logpath": st["n_logpath"], "n_bound": st["n_bound"],
                        "n_noconv": st["n_noconv"], "n_nonpos": st["n_nonpos"],
                        "n_ibb_links": st["n_ibb_links"], "n_repair": st["n_repair"]},
    }
    if args.workload == "c2" and not args.no_learned:
        try:
            out["learned_alpha"] = learned_alpha(pkg, s, L, nu)
        except pkg.UclbsError as exc:  # e.g. a non-finite training loss
            out["learned_alpha"] = {"status": "failed", "error": str(exc)}
    if not args.no_cpu:
        out["cpu_baseline"] = (cpu_baseline_c2(s, L, nu, args.cpu_budget) if args.workload == "c2"
                               else cpu_baseline(args.cpu_budget))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
