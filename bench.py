"""Throughput of the explicit internal-force step on one MI355X.

Workload (one step = one central-difference step of the whole mesh with the
force evaluation inside it): the large Tet4 block — 100^3 hexes split into 6
Tet4 each (6,000,000 elements, 1,030,301 nodes), nodes jittered by 10% of the
edge length, Gaussian initial velocity (sigma 0.5 m/s), compressible
neo-Hookean FP64 (rho 7850, E 200 GPa, nu 0.3), dt = 0.5 x a CFL estimate.
Inputs are resident in HBM before the timed region.

Prints one JSON line.  `python bench.py --gpus 1 --steps K --warmup W`.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "aperi-mech_amd", "python"))

import numpy as np  # noqa: E402

METRIC = "FP64 element-updates/s + % HBM roofline; batched & adjoint vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per element-update of the full-output step: element streams
# conn 16 B + H 72 B + P 72 B; node streams 200 B per node (read X, u, v, a as
# 4 x 24 B and the mass 8 B; write u, v, a, f as 4 x 24 B).
ELEM_BYTES_TET4 = 4 * 4 + 72 + 72
NODE_BYTES = 200
# HBM bytes per step measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (one run each),
# scaled by the same counters on a 1 GiB device copy (fetch x 2.0, write x 1.0); summed
# over the step's two launches.  profiles/tile_step_tet4_6M_neohookean_hbm_traffic.txt
# holds the commands and per-kernel figures.  Keyed by (material, hexes per edge).
MEASURED_TRAFFIC = {("nh", 100): 1399.3e6}
# FP64 flops per step of the same workload from rocprofv3 --pmc SQ_INSTS_VALU_{FMA,MUL,ADD,
# TRANS}_F64 on the tile kernel (64 lanes x (2 FMA + MUL + ADD + TRANS) per wave
# instruction) on the tile and shared-node kernels; profiles/tile_step_tet4_6M_neohookean_
# fp64_valu.txt.  FP64 vector peak of the MI355X: 78.6 TFLOP/s.
MEASURED_FP64_FLOPS = {("nh", 100): 3.32e9}
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(n):
    from aperi_hip import meshgen as M
    return M.large_tet_block(n)


def run_gpu(w, kind, steps, warmup):
    import aperi_hip as A
    from aperi_hip import meshgen as M
    X, conn = w["X"], w["conn"]
    dt = 0.5 * M.stable_dt_estimate(X, conn, w["E"], w["nu"], w["rho"])
    s = A.Solver(device=0)
    t0 = time.time()
    s.set_mesh(A.TET4, X, conn)
    s.set_material(kind, w["rho"], w["E"], w["nu"])
    s.compute_mass()
    s.set_field(A.FIELD_V, w["v0"])
    s.initialize()
    log(f"[bench] set-up {time.time() - t0:.1f}s (nn={len(X)}, ne={len(conn)}, dt={dt:.3e})")
    s.step(dt, warmup)
    s.synchronize()
    s.set_profiling(True)
    s.reset_profile()
    t0 = time.perf_counter()
    s.step(dt, steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    n_step, ms_step = s.profile(0)
    n_force, ms_force = s.profile(1)
    n_node, ms_node = s.profile(2)
    s.set_profiling(False)
    # wall-clock passes without per-kernel events: eager launches, then HIP-graph replay
    walls = {}
    for mode in ("eager", "graph"):
        s.set_graph(mode == "graph")
        s.step(dt, 2)  # (graph: capture and first replay outside the timed region)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(dt, steps)
        s.synchronize()
        walls[mode] = time.perf_counter() - t0
    s.set_graph(False)
    u = s.get(A.FIELD_U)
    ok = bool(np.isfinite(u).all())
    s.close()
    best = min(walls, key=walls.get)
    return dict(wall=walls[best], mode=best, walls=walls, wall_prof=wall, step_ms=ms_step / max(n_step, 1),
                force_ms=ms_force / max(n_force, 1), node_ms=ms_node / max(n_node, 1),
                n_step=n_step, finite=ok, dt=dt)


def bench_batched(n_specimens=256, steps=50):
    """Uneven Hex8 ensemble (edges log-uniform in [8, 48], E ~ U[50, 250] GPa, nu ~ U[0.2, 0.45],
    own dt per specimen): one batched context stepping all specimens (two launches per step,
    replayed from a HIP graph) against the plain loop, one solver per specimen run to
    completion in turn.  Every field of every specimen must be bitwise identical."""
    import aperi_hip as A
    from aperi_hip import meshgen as M
    t0 = time.time()
    specs = M.ensemble(n_specimens)
    dts = [0.5 * M.stable_dt_estimate(sp["X"], sp["conn"], sp["E"], sp["nu"], sp["rho"]) for sp in specs]
    ne = sum(len(sp["conn"]) for sp in specs)
    g = A.Solver()
    g.batch_begin(A.HEX8, A.NEO_HOOKEAN)
    for sp, dt in zip(specs, dts):
        g.batch_add(sp["X"], sp["conn"], sp["rho"], sp["E"], sp["nu"], dt)
    g.batch_end()
    g.compute_mass()
    g.set_field(A.FIELD_V, np.concatenate([sp["v0"] for sp in specs]))
    g.initialize()
    plain = []
    for sp in specs:
        s = A.Solver()
        s.set_mesh(A.HEX8, sp["X"], sp["conn"])
        s.set_material(A.NEO_HOOKEAN, sp["rho"], sp["E"], sp["nu"])
        s.compute_mass()
        s.set_field(A.FIELD_V, sp["v0"])
        s.initialize()
        plain.append(s)
    for s in plain:
        s.synchronize()
    log(f"[bench] ensemble set-up {time.time() - t0:.1f}s ({n_specimens} specimens, {ne} Hex8)")
    g.set_graph(True)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(0.0, steps)  # the increment of each node is its specimen's dt
    g.synchronize()
    t_batch = time.perf_counter() - t0
    t0 = time.perf_counter()
    for s, dt in zip(plain, dts):
        s.step(dt, steps)
        s.synchronize()
    t_loop = time.perf_counter() - t0
    identical = True
    fields = (A.FIELD_U, A.FIELD_V, A.FIELD_A, A.FIELD_F, A.FIELD_H, A.FIELD_P)
    got = {f: g.get(f) for f in fields}
    for k, s in enumerate(plain):
        n0, nn, e0, nel = g.batch_ranges(k)
        for f in fields:
            lo, hi = (e0, e0 + nel) if f in (A.FIELD_H, A.FIELD_P) else (n0, n0 + nn)
            if not np.array_equal(got[f][lo:hi], s.get(f)):
                identical = False
        s.close()
    g.close()
    return {"workload": f"{n_specimens} Hex8 cubes, edges log-uniform in [8,48] ({ne} Hex8), neo-Hookean, "
                        f"per-specimen E, nu and dt, {steps} steps",
            "elements": ne, "steps": steps,
            "batched_s": t_batch, "plain_loop_s": t_loop, "speedup": t_loop / t_batch,
            "batched_element_updates_per_s": ne * steps / t_batch,
            "plain_loop_element_updates_per_s": ne * steps / t_loop,
            "identical_results": identical, "launch": "batched: HIP graph, two launches per step for all "
                                                       "specimens; plain: eager, one solver after another"}


def bench_adjoint(n=64, steps=200, interval=20, n_check=32, h=1e-6):
    """Gradient of L = 1/2 |u_T - u*|^2 with respect to every element's E and nu and the
    initial velocity (64^3 Hex8, smooth E/nu fields), 200 steps with checkpoints every 20,
    against the plain forward; 32 random parameters checked by central differences of the
    same GPU forward (h relative), with L+ - L- formed from the two displacement fields."""
    import aperi_hip as A
    from aperi_hip import meshgen as M
    w = M.adjoint_block(n)
    X, conn, rho = w["X"], w["conn"], w["rho"]
    dt = 0.5 * M.stable_dt_estimate(X, conn, float(w["E"].max()), float(w["nu"].min()), rho)
    s = A.Solver()
    s.set_mesh(A.HEX8, X, conn)
    s.set_material(A.NEO_HOOKEAN, rho, 200e9, 0.3)
    s.compute_mass()

    def forward(E, nu, v0):
        s.set_material_fields(E, nu)
        s.set_field(A.FIELD_U, np.zeros_like(X))
        s.set_field(A.FIELD_V, v0)
        s.initialize()
        s.step(dt, steps)
        return s.get(A.FIELD_U)

    u_star = forward(w["E_target"], w["nu"], w["v0"])
    # plain forward timing
    s.set_material_fields(w["E"], w["nu"])
    s.set_field(A.FIELD_U, np.zeros_like(X))
    s.set_field(A.FIELD_V, w["v0"])
    s.initialize()
    s.synchronize()
    t0 = time.perf_counter()
    s.step(dt, steps)
    s.synchronize()
    t_plain = time.perf_counter() - t0
    # forward with checkpoints + reverse sweep
    s.set_field(A.FIELD_U, np.zeros_like(X))
    s.set_field(A.FIELD_V, w["v0"])
    s.initialize()
    s.synchronize()
    t0 = time.perf_counter()
    s.adjoint_forward(dt, steps, interval)
    s.synchronize()
    t_fwd = time.perf_counter() - t0
    u_T = s.get(A.FIELD_U)
    t0 = time.perf_counter()
    Eb, nub, v0b = s.adjoint_backward(u_T - u_star)
    t_bwd = time.perf_counter() - t0
    # central differences on random parameters (of 32: 12 E, 10 nu, 10 initial-velocity components)
    rng = np.random.default_rng(20261202)
    n_e, n_nu = (12 * n_check) // 32, (10 * n_check) // 32
    picks = ([("E", int(i)) for i in rng.choice(len(conn), n_e, replace=False)] +
             [("nu", int(i)) for i in rng.choice(len(conn), n_nu, replace=False)] +
             [("v0", int(i)) for i in rng.choice(X.size, n_check - n_e - n_nu, replace=False)])
    def central_difference(kind, i, hrel):
        E, nu, v0 = w["E"].copy(), w["nu"].copy(), w["v0"].copy()
        arr = {"E": E, "nu": nu, "v0": v0.reshape(-1)}[kind]
        base = arr[i]
        step = hrel * abs(base)
        arr[i] = base + step
        up = forward(E, nu, v0)
        arr[i] = base - step
        um = forward(E, nu, v0)
        return np.sum((up - um) * (0.5 * (up + um) - u_star)) / (2.0 * step)

    errs, errs10, per_kind = [], [], {}
    for kind, i in picks:
        ad = {"E": Eb, "nu": nub, "v0": v0b.reshape(-1)}[kind][i]
        fd = central_difference(kind, i, h)
        fd10 = central_difference(kind, i, 10 * h)
        errs.append(abs(ad - fd) / max(abs(fd), 1e-300))
        errs10.append(abs(ad - fd10) / max(abs(fd10), 1e-300))
        per_kind.setdefault(kind, []).append(errs[-1])
    s.close()
    errs, errs10 = np.array(errs), np.array(errs10)
    return {"workload": f"{n}^3 Hex8 ({len(conn)} el), neo-Hookean with per-element E/nu fields, "
                        f"{steps} steps, checkpoints every {interval}",
            "plain_forward_s": t_plain, "checkpointed_forward_s": t_fwd, "backward_s": t_bwd,
            "gradient_over_forward": (t_fwd + t_bwd) / t_plain,
            "element_updates_per_s_forward": len(conn) * steps / t_plain,
            "fd_check": {"params": len(picks), "h_rel": h, "max_rel_err": float(errs.max()),
                         "median_rel_err": float(np.median(errs)), "within_1e-6": int((errs < 1e-6).sum()),
                         "max_rel_err_by_kind": {k: float(max(v)) for k, v in per_kind.items()},
                         "max_rel_err_h10": float(errs10.max()), "within_1e-6_h10": int((errs10 < 1e-6).sum()),
                         "reference": "central differences of the same GPU forward (oracle parity of that "
                                      "forward is tested at 1e-12 per step)"}}


def bench_rkpm(n_nodes_edge=64, steps=100, scale=1.5, one_pass=True):
    """Strain-smoothed RKPM: 64^3-node jittered lattice, Kuhn Tet4 background, element-as-cell
    smoothing, kernel support 1.5 x the max edge length, J2 with linear hardening, a velocity
    ramp about the centre reaching a nominal strain of 2 sigma_y / E at step 100.
    one_pass=False: the two-pass force (gather Phi d, element-node cells, scatter Phi^T f)."""
    import aperi_hip as A
    from aperi_hip import meshgen as M
    w = M.rkpm_lattice(n_nodes_edge)
    X, conn = w["X"], w["conn"]
    s = A.Solver()
    s.set_mesh(A.TET4, X, conn)
    s.set_material(A.J2, w["rho"], w["E"], w["nu"], w["yield_stress"], w["hardening"])
    s.synchronize()
    t0 = time.perf_counter()
    s.rk_setup(scale, one_pass=one_pass)
    s.synchronize()
    t_setup = time.perf_counter() - t0
    num, _, truncated = s.rk_neighbors()
    nnz = s.rk_nnz()
    ne, nn = len(conn), len(X)
    s.compute_mass()
    dt = 0.5 * M.stable_dt_estimate(X, conn, w["E"], w["nu"], w["rho"])
    vstar = 2.0 * w["yield_stress"] / w["E"] / (steps * dt)
    s.set_field(A.FIELD_V, M.rkpm_velocity(X, vstar, w["ramp_scale"]))
    s.initialize()
    s.set_profiling(True)
    s.reset_profile()
    t0 = time.perf_counter()
    s.step(dt, steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    n0, m0 = s.profile(0)
    n1, m1 = s.profile(1)
    n2, m2 = s.profile(2)
    yielded = int((s.get(A.FIELD_STATE) > 0).sum())
    s.close()
    lc = nnz / ne
    # one-pass, per cell: entries (4 B id + 24 B grad) + H, P, volume, J2 state N/NP1;
    # per node: the node streams.  Two-pass, per cell: the four node ids + H, P, J2 state
    # (the derivatives and volume follow from X, part of the node streams); per node: the
    # node streams, the neighbor rows of the gather and of the scatter (4 B id + 8 B phi
    # per neighbor each) and the physical displacement and local force, each written
    # once and read once
    if one_pass:
        bytes_per_update = 28.0 * lc + 168.0 + 200.0 * nn / ne
    else:
        bytes_per_update = 16.0 + 160.0 + (200.0 + 96.0 + 24.0 * float(num.mean())) * nn / ne
    step_ms = m0 / max(n0, 1)
    achieved = bytes_per_update * ne / (step_ms * 1e-3) / 1e9
    method = "one-pass" if one_pass else "two-pass"
    return {"workload": f"{n_nodes_edge}^3-node RKPM lattice ({nn} nodes, {ne} Tet4 cells), J2, kernel scale {scale}, "
                        f"{method} force, {steps} steps",
            "setup_s": t_setup, "mean_neighbors": float(num.mean()), "max_neighbors": int(num.max()),
            "truncated_nodes": int(truncated), "cell_list_mean": lc,
            "cell_updates_per_s": ne * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "kernels_ms": {"step": step_ms, "cell_force" if one_pass else "gather_and_cell_force": m1 / max(n1, 1),
                           "node_gather" if one_pass else "node_sum_and_scatter": m2 / max(n2, 1)},
            "roofline": {"bound": "hbm", "bytes_per_update": bytes_per_update, "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS},
            "yielded_cells": yielded}


def cpu_baseline(w, kind, budget_s=12.0):
    """Oracle (C restatement of the reference data flow) on all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    X, conn = w["X"], w["conn"]
    lam, mu = O.lame(w["E"], w["nu"])
    mass = O.lumped_mass(0, X, conn, w["rho"])
    from aperi_hip import meshgen as M
    dt = 0.5 * M.stable_dt_estimate(X, conn, w["E"], w["nu"], w["rho"])
    threads = min(os.cpu_count() or 1, 16)
    ref = O.Solver(0, X, conn, kind, lam, mu, mass)
    ref.v[:] = w["v0"]
    ref.run(dt, 0, nthreads=threads)
    t0 = time.perf_counter()
    ref.run(dt, 1, nthreads=threads)
    one = time.perf_counter() - t0
    n = int(max(1, min(20, budget_s / max(one, 1e-3))))
    t0 = time.perf_counter()
    ref.run(dt, n, nthreads=threads)
    el = time.perf_counter() - t0
    ne = conn.shape[0]
    t0 = time.perf_counter()
    ref.run(dt, 1, nthreads=1)
    one_thread = time.perf_counter() - t0
    return dict(value=ne * n / el, unit="element-updates/s", cores=threads, kind="port",
                sample=f"{n} steps of the same 6M-tet neo-Hookean workload after 1 warm step "
                       f"({el:.1f}s, OpenMP atomic scatter, oracle/liboracle.so)",
                single_thread_value=ne / one_thread)


def fp64_line(flops_per_step, step_ms):
    """FP64 VALU use beside the HBM roofline: counter-measured flops per step over the
    event-timed step."""
    if flops_per_step is None:
        return None
    achieved = flops_per_step / (step_ms * 1e-3) / 1e12
    return {"flops_per_step": flops_per_step, "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS}


def bench_power(n=100):
    """Power-method stable time increment (the reference's PowerMethodProcessor algorithm)
    on the large Tet4 block at the reference's default settings: each iteration is one
    perturbed internal-force evaluation plus a norm; reports the wall time, the iteration
    count, the increment and its ratio to the element-size estimate used by the step."""
    import aperi_hip as A
    from aperi_hip import meshgen as M
    w = workload(n)
    X, conn = w["X"], w["conn"]
    s = A.Solver(device=0)
    s.set_mesh(A.TET4, X, conn)
    s.set_material(A.NEO_HOOKEAN, w["rho"], w["E"], w["nu"])
    s.compute_mass()
    s.set_field(A.FIELD_V, w["v0"])
    s.initialize()
    est = M.stable_dt_estimate(X, conn, w["E"], w["nu"], w["rho"])
    s.step(0.5 * est, 2)  # a deformed state: the increment is evaluated about the current displacement
    s.synchronize()
    s.power_method_time_increment(num_iterations=3)  # warm-up (allocations)
    t0 = time.perf_counter()
    dt, it, cv = s.power_method_time_increment()
    s.synchronize()
    el = time.perf_counter() - t0
    s.close()
    return {"wall_s": el, "iterations": it, "converged": cv, "ms_per_iteration": 1e3 * el / max(it, 1),
            "time_increment": dt, "over_element_estimate": dt / est,
            "workload": f"{len(conn)} Tet4, {len(X)} nodes, neo-Hookean, 50 iterations max, tolerance 1e-3"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=100, help="hexes per edge (x6 tets)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--material", default="nh", choices=["nh", "le"])
    ap.add_argument("--no-extras", action="store_true", help="skip the ensemble, RKPM, power-method and adjoint measurements")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only time a device copy of BYTES (HBM yardstick) and print it")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(the force step would shard by Morton tile ranges with a shared-node force exchange)",
              file=sys.stderr)
        sys.exit(2)
    import aperi_hip as A
    if args.copy_probe:
        ms = A.bandwidth_probe(args.copy_probe, reps=max(args.steps, 1))
        print(json.dumps({"copy_bytes": args.copy_probe, "ms_per_copy": ms,
                          "GBps": 2 * args.copy_probe / (ms * 1e-3) / 1e9}), flush=True)
        return
    kind = A.NEO_HOOKEAN if args.material == "nh" else A.LINEAR_ELASTIC
    t0 = time.time()
    w = workload(args.n)
    log(f"[bench] workload generated in {time.time() - t0:.1f}s")
    r = run_gpu(w, kind, args.steps, args.warmup)
    ne, nn = w["conn"].shape[0], w["X"].shape[0]
    value = ne * args.steps / r["wall"]
    bytes_per_update = ELEM_BYTES_TET4 + NODE_BYTES * nn / ne
    achieved = bytes_per_update * ne / (r["step_ms"] * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "element-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered lattice, counter-based splitmix64)",
        "config": {"workload": f"large Tet4 block {args.n}^3x6 = {ne} Tet4, {nn} nodes, "
                               f"{'neo-Hookean' if kind == A.NEO_HOOKEAN else 'linear elastic'} FP64, "
                               f"full-output central-difference step",
                   "elements": ne, "nodes": nn, "material": args.material},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC.get((args.material, args.n)),
                     "kernel": "one step = k_tile (fused element tile) + k_boundary (shared nodes), "
                               "one launch each; achieved = algorithmic bytes / event-timed step",
                     "bytes_per_update": bytes_per_update,
                     "algorithmic_bytes_per_step": bytes_per_update * ne},
        "kernels_ms": {"step": r["step_ms"], "tile_kernel": r["force_ms"], "shared_node_kernel": r["node_ms"]},
        "fp64_valu": fp64_line(MEASURED_FP64_FLOPS.get((args.material, args.n)), r["step_ms"]),
        "launch": {"mode": r["mode"], "ms_per_step": {k: 1e3 * v / args.steps for k, v in r["walls"].items()},
                   "tiling": "1024 Tet4 per workgroup, one per thread (default plan)"},
    }
    if not r["finite"]:
        line["error"] = "non-finite displacement"
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(w, kind)
    del w
    if not args.no_extras:
        log("[bench] batched ensemble ...")
        line["batched"] = bench_batched()
        log(f"[bench] batched: {line['batched']['speedup']:.2f}x, identical={line['batched']['identical_results']}")
        log("[bench] rkpm ...")
        line["rkpm"] = bench_rkpm()
        log(f"[bench] rkpm: {line['rkpm']['cell_updates_per_s']:.3e} cell-updates/s, "
            f"setup {line['rkpm']['setup_s']:.2f}s, roofline {line['rkpm']['roofline']['frac']:.3f}")
        log("[bench] rkpm two-pass ...")
        line["rkpm_two_pass"] = bench_rkpm(one_pass=False)
        log(f"[bench] rkpm two-pass: {line['rkpm_two_pass']['cell_updates_per_s']:.3e} cell-updates/s, "
            f"roofline {line['rkpm_two_pass']['roofline']['frac']:.3f}")
        log("[bench] power method ...")
        line["power_method"] = bench_power(args.n)
        log(f"[bench] power method: {line['power_method']['iterations']} iterations, "
            f"{line['power_method']['ms_per_iteration']:.2f} ms each")
        log("[bench] adjoint ...")
        line["adjoint"] = bench_adjoint()
        log(f"[bench] adjoint: {line['adjoint']['gradient_over_forward']:.2f}x forward, "
            f"fd max rel err {line['adjoint']['fd_check']['max_rel_err']:.2e}")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
