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
MEASURED_TRAFFIC = {("nh", 100): 1339.7e6, ("le", 100): 1337.9e6}
# FP64 flops per step of the same workload from rocprofv3 --pmc SQ_INSTS_VALU_{FMA,MUL,ADD,
# TRANS}_F64 on the tile kernel (64 lanes x (2 FMA + MUL + ADD + TRANS) per wave
# instruction) on the tile and shared-node kernels; profiles/tile_step_tet4_6M_neohookean_
# fp64_valu.txt.  FP64 vector peak of the MI355X: 78.6 TFLOP/s.
MEASURED_FP64_FLOPS = {("nh", 100): 3.30e9, ("le", 100): 1.64e9}
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


def hex8_bytes_per_update(nn, ne, material_fields=True):
    """Algorithmic bytes per Hex8 element-update: connectivity 32 B, H 72 B, P 72 B, the
    per-element E and nu 16 B when the material varies per element, and the node streams
    (200 B per node) spread over the elements."""
    return 8 * 4 + 72 + 72 + (16 if material_fields else 0) + NODE_BYTES * nn / ne


def roofline(bytes_per_update, updates_per_launch, ms_per_launch, kernel):
    achieved = bytes_per_update * updates_per_launch / (ms_per_launch * 1e-3) / 1e9
    return {"bound": "hbm", "bytes_per_update": bytes_per_update, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "kernel": kernel}


def bench_batched(n_specimens=256, steps=50, tune=True):
    """Uneven Hex8 ensemble (edges log-uniform in [8, 48], E ~ U[50, 250] GPa, nu ~ U[0.2, 0.45],
    own dt per specimen): one batched context stepping all specimens (two launches per step,
    replayed from a HIP graph) against the plain loop, one solver per specimen run to
    completion in turn.  With the default tile plan on both sides every field of every
    specimen must be bitwise identical.  Then both sides are autotuned (the batched context
    once, every plain specimen on its own) and timed again over the same number of steps;
    different plans change the summation grouping, so those results agree to rounding."""
    import aperi_hip as A
    from aperi_hip import meshgen as M
    t0 = time.time()
    specs = M.ensemble(n_specimens)
    dts = [0.5 * M.stable_dt_estimate(sp["X"], sp["conn"], sp["E"], sp["nu"], sp["rho"]) for sp in specs]
    ne = sum(len(sp["conn"]) for sp in specs)
    nn = sum(len(sp["X"]) for sp in specs)
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

    def run_batched():
        g.set_graph(True)
        g.synchronize()
        t0 = time.perf_counter()
        g.step(0.0, steps)  # the increment of each node is its specimen's dt
        g.synchronize()
        return time.perf_counter() - t0

    def run_loop():
        t0 = time.perf_counter()
        for s, dt in zip(plain, dts):
            s.step(dt, steps)
            s.synchronize()
        return time.perf_counter() - t0

    def kernel_ms():
        # per-step device time of the batched step from HIP events on the context's stream
        g.set_graph(False)
        g.set_profiling(True)
        g.reset_profile()
        g.step(0.0, 10)
        g.synchronize()
        n0, m0 = g.profile(0)
        n1, m1 = g.profile(1)
        g.set_profiling(False)
        for s, dt in zip(plain, dts):  # keep the loop in step with the batched context
            s.step(dt, 10)
        return m0 / max(n0, 1), m1 / max(n1, 1)

    fields = (A.FIELD_U, A.FIELD_V, A.FIELD_A, A.FIELD_F, A.FIELD_H, A.FIELD_P)

    def compare(bitwise):
        got = {f: g.get(f) for f in fields}
        same, worst = True, 0.0
        for k, s in enumerate(plain):
            n0, nn_k, e0, nel = g.batch_ranges(k)
            for f in fields:
                lo, hi = (e0, e0 + nel) if f in (A.FIELD_H, A.FIELD_P) else (n0, n0 + nn_k)
                a, b = got[f][lo:hi], s.get(f)
                same = same and np.array_equal(a, b)
                worst = max(worst, float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)))
        return same, worst

    t_batch = run_batched()
    t_loop = run_loop()
    identical, _ = compare(True)
    step_ms, tile_ms = kernel_ms()
    bpu = hex8_bytes_per_update(nn, ne, material_fields=True)
    out = {"workload": f"{n_specimens} Hex8 cubes, edges log-uniform in [8,48] ({ne} Hex8, {nn} nodes), "
                       f"neo-Hookean, per-specimen E, nu and dt, {steps} steps",
           "elements": ne, "nodes": nn, "steps": steps,
           "batched_s": t_batch, "plain_loop_s": t_loop, "speedup": t_loop / t_batch,
           "batched_element_updates_per_s": ne * steps / t_batch,
           "plain_loop_element_updates_per_s": ne * steps / t_loop,
           "identical_results": identical,
           "kernels_ms": {"step": step_ms, "tile_kernel": tile_ms},
           "roofline": roofline(bpu, ne, step_ms, "batched step (tile kernel + shared-node kernel), HIP events"),
           "launch": "batched: HIP graph, two launches per step for all specimens; plain: eager, one solver "
                     "after another; default tile plan (128 Hex8 per workgroup) on both sides"}
    if tune:
        t0 = time.time()
        choice = g.autotune(0.0, 3)
        loop_choices = {}
        for s, dt in zip(plain, dts):
            c = s.autotune(dt, 3)
            loop_choices[c] = loop_choices.get(c, 0) + 1
        log(f"[bench] ensemble autotune {time.time() - t0:.1f}s (batched plan {choice})")
        tb = run_batched()
        tl = run_loop()
        _, worst = compare(False)
        out["autotuned"] = {"batched_s": tb, "plain_loop_s": tl, "speedup": tl / tb,
                            "batched_plan": choice,
                            "plain_loop_plans": {str(k): v for k, v in sorted(loop_choices.items())},
                            "max_rel_diff_after_further_steps": worst,
                            "plan_code": "elements_per_tile * 10000 + threads"}
        # the batched context's tuned plan on every plain specimen too, each specimen
        # restarted from the batched context's state: the same tiles on both sides, hence
        # bitwise identical results again
        te, thr = divmod(choice % 100000000, 10000)
        gu, gv, ga = g.get(A.FIELD_U), g.get(A.FIELD_V), g.get(A.FIELD_A)
        for k, s in enumerate(plain):
            n0, nn_k, _, _ = g.batch_ranges(k)
            s.set_tiling(te, thr)
            s.set_field(A.FIELD_U, gu[n0:n0 + nn_k])
            s.set_field(A.FIELD_V, gv[n0:n0 + nn_k])
            s.set_field(A.FIELD_A, ga[n0:n0 + nn_k])
        tb2 = run_batched()
        tl2 = run_loop()
        same2, _ = compare(True)
        out["same_tuned_plan"] = {"plan": choice, "batched_s": tb2, "plain_loop_s": tl2, "speedup": tl2 / tb2,
                                  "identical_results": same2,
                                  "note": "the batched context's autotuned plan set on every plain specimen"}
    for s in plain:
        s.close()
    g.close()
    return out


def bench_adjoint(n=64, steps=200, interval=20, n_check=32, h=1e-6):
    """Gradient of L = 1/2 |u_T - u*|^2 with respect to every element's E and nu and the
    initial velocity (64^3 Hex8, smooth E/nu fields), 200 steps with checkpoints every 20,
    against the plain forward; 32 random parameters checked by central differences of the
    same GPU forward (whose parity with the oracle over the 200 steps is a GPU test), with
    L+ - L- formed from the two displacement fields.  Steps: h |theta| for E_e and nu_e;
    an initial-velocity component, which can sit arbitrarily close to zero, takes 10 h
    times the velocity field's amplitude (at h its smallest sensitivities sit at the
    rounding floor of two 200-step forwards); the figures with h times the amplitude and
    with h relative to each component's own value are reported beside it."""
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
    # per-step device time of the forward step (HIP events, eager)
    s.set_profiling(True)
    s.reset_profile()
    s.step(dt, 20)
    s.synchronize()
    nfw, mfw = s.profile(0)
    s.set_profiling(False)
    fwd_ms = mfw / max(nfw, 1)
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
    vscale = float(np.abs(w["v0"]).max())

    def central_difference(kind, i, step):
        E, nu, v0 = w["E"].copy(), w["nu"].copy(), w["v0"].copy()
        arr = {"E": E, "nu": nu, "v0": v0.reshape(-1)}[kind]
        base = arr[i]
        arr[i] = base + step
        up = forward(E, nu, v0)
        arr[i] = base - step
        um = forward(E, nu, v0)
        return np.sum((up - um) * (0.5 * (up + um) - u_star)) / (2.0 * step)

    errs, errs_own, errs_h, errs10, per_kind = [], [], [], [], {}
    for kind, i in picks:
        ad = {"E": Eb, "nu": nub, "v0": v0b.reshape(-1)}[kind][i]
        theta = {"E": w["E"], "nu": w["nu"], "v0": w["v0"].reshape(-1)}[kind][i]
        # the documented step per kind: h |theta| for E and nu, 10 h x the velocity amplitude
        # for initial-velocity components (whose smallest sensitivities sit at the rounding
        # floor of the 200-step forward with h)
        step = h * abs(theta) if kind in ("E", "nu") else 10 * h * vscale
        fd = central_difference(kind, i, step)
        errs.append(abs(ad - fd) / max(abs(fd), 1e-300))
        per_kind.setdefault(kind, []).append(errs[-1])
        if kind == "v0":
            fd_h = central_difference(kind, i, h * vscale)
            errs_h.append(abs(ad - fd_h) / max(abs(fd_h), 1e-300))
            fd_own = central_difference(kind, i, h * abs(theta))
            errs_own.append(abs(ad - fd_own) / max(abs(fd_own), 1e-300))
        base = abs(theta) if kind in ("E", "nu") else vscale
        fd10 = central_difference(kind, i, 10 * h * base)
        errs10.append(abs(ad - fd10) / max(abs(fd10), 1e-300))
    s.close()
    errs, errs10 = np.array(errs), np.array(errs10)
    ne, nn = len(conn), len(X)
    bpu_fwd = hex8_bytes_per_update(nn, ne, material_fields=True)
    # reverse sweep per element-update: the forward's streams again (segment recompute), the
    # adjoint node vectors lambda_u and lambda_v read and written (4 x 24 B per node) and the
    # per-element E / nu sensitivities read and written (32 B)
    bpu_rev = bpu_fwd + 4 * 24 * nn / ne + 32
    rev_ms = 1e3 * t_bwd / steps
    return {"workload": f"{n}^3 Hex8 ({ne} el), neo-Hookean with per-element E/nu fields, "
                        f"{steps} steps, checkpoints every {interval}",
            "plain_forward_s": t_plain, "checkpointed_forward_s": t_fwd, "backward_s": t_bwd,
            "gradient_over_forward": (t_fwd + t_bwd) / t_plain,
            "element_updates_per_s_forward": ne * steps / t_plain,
            "kernels_ms": {"forward_step": fwd_ms, "reverse_step_wall": rev_ms},
            "roofline": roofline(bpu_fwd, ne, fwd_ms, "forward step (tile kernel + shared-node kernel), HIP events"),
            "roofline_reverse": roofline(bpu_rev, ne, rev_ms,
                                         "reverse sweep per step incl. its segment recompute, wall clock"),
            "fd_check": {"params": len(picks), "h_rel": h,
                         "step": "h x |theta| for E_e and nu_e, 10 h x max|v0| for initial-velocity components",
                         "max_rel_err": float(errs.max()),
                         "median_rel_err": float(np.median(errs)), "within_1e-6": int((errs < 1e-6).sum()),
                         "max_rel_err_by_kind": {k: float(max(v)) for k, v in per_kind.items()},
                         "v0_step_h_times_amplitude": {"max_rel_err": float(max(errs_h)),
                                                       "within_1e-6": int((np.array(errs_h) < 1e-6).sum()),
                                                       "params": len(errs_h)},
                         "v0_step_relative_to_component": {"max_rel_err": float(max(errs_own)),
                                                           "within_1e-6": int((np.array(errs_own) < 1e-6).sum()),
                                                           "params": len(errs_own)},
                         "max_rel_err_h10": float(errs10.max()), "within_1e-6_h10": int((errs10 < 1e-6).sum()),
                         "reference": "central differences of the same GPU forward (its parity with the oracle "
                                      "over these 200 steps is tested at 1e-10)"}}


def bench_hex8_forward(n=64, steps=20):
    """The plain forward of the adjoint block alone (64^3 Hex8, neo-Hookean with per-element
    E / nu fields): event-timed step and its roofline; used for kernel traces and counters."""
    import aperi_hip as A
    from aperi_hip import meshgen as M
    w = M.adjoint_block(n)
    X, conn, rho = w["X"], w["conn"], w["rho"]
    dt = 0.5 * M.stable_dt_estimate(X, conn, float(w["E"].max()), float(w["nu"].min()), rho)
    s = A.Solver()
    s.set_mesh(A.HEX8, X, conn)
    s.set_material(A.NEO_HOOKEAN, rho, 200e9, 0.3)
    s.set_material_fields(w["E"], w["nu"])
    s.compute_mass()
    s.set_field(A.FIELD_V, w["v0"])
    s.initialize()
    s.step(dt, 2)
    s.set_profiling(True)
    s.reset_profile()
    s.step(dt, steps)
    s.synchronize()
    n0, m0 = s.profile(0)
    n1, m1 = s.profile(1)
    s.close()
    ne, nn = len(conn), len(X)
    step_ms = m0 / max(n0, 1)
    return {"workload": f"{n}^3 Hex8 ({ne} el), neo-Hookean with per-element E/nu, {steps} steps",
            "kernels_ms": {"step": step_ms, "tile_kernel": m1 / max(n1, 1)},
            "roofline": roofline(hex8_bytes_per_update(nn, ne, True), ne, step_ms, "step, HIP events")}


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
            "extra_neighbors": int(truncated), "cell_list_mean": lc,
            "cell_updates_per_s": ne * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "kernels_ms": {"step": step_ms, "cell_force" if one_pass else "gather_and_cell_force": m1 / max(n1, 1),
                           "node_gather" if one_pass else "node_sum_and_scatter": m2 / max(n2, 1)},
            "roofline": {"bound": "hbm", "bytes_per_update": bytes_per_update, "achieved": achieved,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS},
            "yielded_cells": yielded}


def _affinity():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# The process's CPU share, read at start-up: once libgomp loads with OMP_PROC_BIND set it
# binds the calling thread to one place, and the affinity mask read afterwards shrinks.
_CPUS_ALLOWED = _affinity()


def host_cpu():
    """CPU model and the CPUs this process may run on (its affinity mask at start-up)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, _CPUS_ALLOWED


def cpu_threads():
    """Threads for the CPU baseline: every CPU of the affinity mask, capped by
    OMP_NUM_THREADS when the environment sets one (the host's CPU share)."""
    _, allowed = host_cpu()
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(allowed, env) if env > 0 else allowed)


def pin_openmp_threads():
    """Pinned OpenMP threads (one per core, placed close together).  libgomp reads these
    when it is loaded, which happens with the oracle library, after this call."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")


def bench_nodal(n=20, steps=100, f_bar=True):
    """Nodal strain smoothing on a thex mesh, modelled on the reference's Taylor-bar
    performance inputs (test/performance_tests/aperi-mech/taylor_bar/rkpm_fbar and
    rkpm_nodal input.yaml: thex mesh, reproducing kernel, elastic (SVK) material with
    E = 78.2e8 Pa, nu = 0.3, rho = 2700; F-bar: every hex a subcell, kernel scale 0.8;
    nodal: one subcell per cell, scale 0.9): a jittered n^3 hex lattice split into Kuhn
    tets, each tet into four hexes, an impact-like initial velocity of -373 m/s in z."""
    import aperi_hip as A
    from aperi_hip import meshgen as M
    Xh, hx = M.thex_block(n)
    rho, E, nu = 2700.0, 78.2e8, 0.3
    scale, subdomains = (0.8, 0) if f_bar else (0.9, 1)
    s = A.Solver()
    s.set_mesh(A.HEX8, Xh, hx)
    s.set_material(A.SVK, rho, E, nu)
    s.synchronize()
    t0 = time.perf_counter()
    s.rk_nodal_setup(scale, subdomains, f_bar)
    s.synchronize()
    t_setup = time.perf_counter() - t0
    cs, es, act, fb = s.rk_nodal_cells()
    nnz = s.rk_nnz()
    n_sub = int(es.max()) + 1
    n_act = int(act.sum())
    s.compute_mass()
    v0 = np.zeros_like(Xh)
    v0[:, 2] = -373.0
    dt = 0.25 * M.stable_dt_estimate(Xh, hx, E, nu, rho)
    s.set_field(A.FIELD_V, v0)
    s.initialize()
    s.step(dt, 5)
    s.synchronize()
    s.set_profiling(True)
    s.reset_profile()
    t0 = time.perf_counter()
    s.step(dt, steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    n0, m0 = s.profile(0)
    n1, m1 = s.profile(1)
    n2, m2 = s.profile(2)
    ok = bool(np.isfinite(s.get(A.FIELD_U)).all())
    s.close()
    # per subcell-update: entries (4 B id + 24 B grad phi), H and P written (144 B), the
    # volume read (8 B); the node streams (200 B per node with degrees of freedom)
    lc = nnz / n_sub
    bytes_per_update = 28.0 * lc + 152.0 + 200.0 * n_act / n_sub
    step_ms = m0 / max(n0, 1)
    return {"workload": f"thex of a {n}^3 Kuhn-tet lattice ({len(hx)} Hex8, {len(Xh)} nodes, {n_act} active), "
                        f"SVK, kernel scale {scale}, {'F-bar, every hex a subcell' if f_bar else 'one subcell per cell'}, "
                        f"{steps} steps",
            "cells": int(len(cs) - 1), "subcells": n_sub, "f_bar": fb, "setup_s": t_setup,
            "subcell_list_mean": lc, "subcell_updates_per_s": n_sub * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "kernels_ms": {"step": step_ms, "cell_force": m1 / max(n1, 1), "node_gather": m2 / max(n2, 1)},
            "roofline": roofline(bytes_per_update, n_sub, step_ms, "one step (cell passes + node kernel + "
                                                                    "predictor), HIP events"),
            "finite": ok}


def cpu_baseline(w, kind, budget_s=12.0, single_thread=True):
    """Oracle (C restatement of the reference data flow) on all host cores of the process's
    CPU share, threads pinned."""
    pin_openmp_threads()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    X, conn = w["X"], w["conn"]
    lam, mu = O.lame(w["E"], w["nu"])
    mass = O.lumped_mass(0, X, conn, w["rho"])
    from aperi_hip import meshgen as M
    dt = 0.5 * M.stable_dt_estimate(X, conn, w["E"], w["nu"], w["rho"])
    threads = cpu_threads()
    model, allowed = host_cpu()
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
    out = dict(value=ne * n / el, unit="element-updates/s", cores=threads, kind="port",
               sample=f"{n} steps of the same {ne}-Tet4 {'neo-Hookean' if kind == 1 else 'linear-elastic'} workload "
                      f"after 1 warm step ({el:.1f}s, OpenMP atomic scatter, oracle/liboracle.so, run in a "
                      f"child process of the bench)",
               cpu_model=model, cpus_allowed=allowed, cpus_online=os.cpu_count(),
               omp_proc_bind=os.environ.get("OMP_PROC_BIND"), omp_places=os.environ.get("OMP_PLACES"))
    if single_thread:
        t0 = time.perf_counter()
        ref.run(dt, 1, nthreads=1)
        out["single_thread_value"] = ne / (time.perf_counter() - t0)
    return out


def cpu_baseline_child(n, material, budget_s, single_thread):
    """cpu_baseline in a fresh interpreter (this file with --cpu-baseline-only): the
    child's start-up affinity is the host's CPU share, and the OpenMP binding stays in it."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", material, "--n", str(n),
           "--cpu-budget", str(budget_s), "--single-thread", "1" if single_thread else "0"]
    out = subprocess.run(cmd, check=True, stdout=subprocess.PIPE, text=True).stdout
    return json.loads(out.strip().splitlines()[-1])


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
    ap.add_argument("--cpu-baseline-only", default=None, choices=["nh", "le"],
                    help="only time the oracle on the block (run by the bench itself in a child process)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--single-thread", type=int, default=1)
    args = ap.parse_args()
    if args.cpu_baseline_only:
        okind = 1 if args.cpu_baseline_only == "nh" else 0
        print(json.dumps(cpu_baseline(workload(args.n), okind, args.cpu_budget, bool(args.single_thread))), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(the force step would shard by Morton tile ranges with a shared-node force exchange)",
              file=sys.stderr)
        sys.exit(2)
    baselines = {}
    if not args.no_cpu and not args.copy_probe:
        # The oracle runs in child processes started before this process touches the GPU:
        # libgomp binds the thread that loads it (OMP_PROC_BIND), which would otherwise pin
        # this process's host work (set-up threads included) to one core.
        log("[bench] cpu baseline ...")
        mats = [args.material] + ([] if args.no_extras else ["le" if args.material == "nh" else "nh"])
        for k, mat in enumerate(mats):
            baselines[mat] = cpu_baseline_child(args.n, mat, 12.0 if k == 0 else 6.0, k == 0)
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
    if args.material in baselines:
        line["cpu_baseline"] = baselines[args.material]
    if not args.no_extras:
        # the other material of the same block: GPU step and CPU baseline
        other = A.LINEAR_ELASTIC if kind == A.NEO_HOOKEAN else A.NEO_HOOKEAN
        oname = "le" if other == A.LINEAR_ELASTIC else "nh"
        log(f"[bench] {oname} variant ...")
        ro = run_gpu(w, other, args.steps, args.warmup)
        leg = {"material": oname, "value": ne * args.steps / ro["wall"], "unit": "element-updates/s",
               "ms_per_step": 1e3 * ro["wall"] / args.steps,
               "kernels_ms": {"step": ro["step_ms"], "tile_kernel": ro["force_ms"], "shared_node_kernel": ro["node_ms"]},
               "roofline": roofline(bytes_per_update, ne, ro["step_ms"],
                                    "one step = k_tile + k_boundary, HIP events"),
               "fp64_valu": fp64_line(MEASURED_FP64_FLOPS.get((oname, args.n)), ro["step_ms"]),
               "traffic": MEASURED_TRAFFIC.get((oname, args.n))}
        if oname in baselines:
            leg["cpu_baseline"] = baselines[oname]
        line[f"{oname}_variant"] = leg
        log(f"[bench] {oname} variant: {leg['value']:.3e} element-updates/s, roofline {leg['roofline']['frac']:.3f}")
    del w
    if not args.no_extras:
        log("[bench] batched ensemble ...")
        line["batched"] = bench_batched()
        log(f"[bench] batched: {line['batched']['speedup']:.2f}x, identical={line['batched']['identical_results']}, "
            f"roofline {line['batched']['roofline']['frac']:.3f}, autotuned "
            f"{line['batched'].get('autotuned', {}).get('speedup', float('nan')):.2f}x, same tuned plan "
            f"{line['batched'].get('same_tuned_plan', {}).get('speedup', float('nan')):.2f}x "
            f"(identical={line['batched'].get('same_tuned_plan', {}).get('identical_results')})")
        log("[bench] rkpm ...")
        line["rkpm"] = bench_rkpm()
        log(f"[bench] rkpm: {line['rkpm']['cell_updates_per_s']:.3e} cell-updates/s, "
            f"setup {line['rkpm']['setup_s']:.2f}s, roofline {line['rkpm']['roofline']['frac']:.3f}")
        log("[bench] rkpm two-pass ...")
        line["rkpm_two_pass"] = bench_rkpm(one_pass=False)
        log(f"[bench] rkpm two-pass: {line['rkpm_two_pass']['cell_updates_per_s']:.3e} cell-updates/s, "
            f"roofline {line['rkpm_two_pass']['roofline']['frac']:.3f}")
        for fb in (True, False):
            key = "nodal_fbar" if fb else "nodal"
            log(f"[bench] {key} ...")
            line[key] = bench_nodal(f_bar=fb)
            log(f"[bench] {key}: {line[key]['subcell_updates_per_s']:.3e} subcell-updates/s, "
                f"setup {line[key]['setup_s']:.2f}s, roofline {line[key]['roofline']['frac']:.3f}")
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
