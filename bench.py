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
MEASURED_TRAFFIC = {("nh", 100): 1451.0e6}


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
    return dict(value=ne * n / el, unit="element-updates/s", cores=threads, kind="port",
                sample=f"{n} steps of the same 6M-tet neo-Hookean workload after 1 warm step "
                       f"({el:.1f}s, OpenMP atomic scatter, oracle/liboracle.so)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=100, help="hexes per edge (x6 tets)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--material", default="nh", choices=["nh", "le"])
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
        "launch": {"mode": r["mode"], "ms_per_step": {k: 1e3 * v / args.steps for k, v in r["walls"].items()},
                   "tiling": "512 Tet4 per workgroup, one per thread (default plan)"},
    }
    if not r["finite"]:
        line["error"] = "non-finite displacement"
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(w, kind)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
