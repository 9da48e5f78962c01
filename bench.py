"""Benchmark: DOF*stage updates per second of the FP64 DG-SEM wave RK3 stage on one MI355X.

Workload (one GPU): 512 x 512 quads on [-1,1]^2, interior vertices jittered by a seeded uniform
+-15% of the spacing, elements in Hilbert-curve order, N = 8 (81 Legendre-Gauss nodes per
element, 63,700,992 DOF), reflecting walls (the reference's wall ghost), Gaussian pulses of
amplitude 1 at (0.1, -0.2) and 0.5 at (-0.4, 0.3), width 0.2. A "step" is one low-storage RK3 step
(3 fused stage kernels) over the whole mesh, inputs resident in HBM.

Prints one JSON line. Timing: HIP events recorded on the stream the kernels run on, around the
K timed steps (after W warm-up steps), bracketed by device synchronisation.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

RES, N, JITTER, SEED = 512, 8, 0.15, 2024
PULSES = [(0.1, -0.2, 0.2, 1.0), (-0.4, 0.3, 0.2, 0.5)]
CFL = 0.5
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stage_kernel_N8_512x512_traffic.json")
TRAFFIC_FILE_C32 = os.path.join(ROOT, "profiles", "stage_c32_kernel_N8_512x512_traffic.json")


def algorithmic_bytes_per_element(n):
    """FP64 bytes of the reference's stage per element, averaged over the 3 stages, for a minimal
    fused implementation of it: state and RK register (stage 1 reads q, writes q and g; stages 2-3
    read and write both) 88*(N+1)^2, neighbour traces read plus own traces written 192*(N+1),
    vertices and neighbour ids 80. The fused kernel here also skips the last stage's RK register
    store (the next step's first stage overwrites it unread), so it draws 8*(N+1)^2 less from HBM;
    roofline.hbm_gbs reports the bandwidth it actually draws (measured traffic / launch time)."""
    nf = n + 1
    return 88 * nf * nf + 192 * nf + 80


def algorithmic_bytes_fp32c(n):
    """The FP32 form per element and stage: state as two floats read and written 48*(N+1)^2, RK
    register as floats 20*(N+1)^2 (2/3 read + written), float traces 96*(N+1), vertices and ids 80."""
    nf = n + 1
    return 68 * nf * nf + 96 * nf + 80


# Mixed-order (batched) workload: 256 x 256 walls, 8 seeded pulses, element orders 4..12 graded by
# distance to the nearest pulse centre (a static snapshot of p-refinement around the pulses).
B_RES, B_SEED = 256, 77


def batched_workload():
    rng = np.random.default_rng(B_SEED)
    pulses = np.column_stack([rng.uniform(-0.7, 0.7, 8), rng.uniform(-0.7, 0.7, 8),
                              rng.uniform(0.05, 0.2, 8), rng.uniform(0.2, 1.0, 8)])
    return pulses


def batched_orders(desc, pulses):
    xy = desc.xy[desc.quads].mean(axis=1)
    d = np.min(np.hypot(xy[:, None, 0] - pulses[None, :, 0], xy[:, None, 1] - pulses[None, :, 1])
               / pulses[None, :, 2], axis=1)
    return np.select([d < 1.0, d < 2.0, d < 3.0, d < 4.5], [12, 10, 8, 6], 4).astype(np.int32)


def mortar_bytes(orders, desc):
    """Bytes the mortar kernel moves per stage: both side traces read and both projected side
    fluxes written (3 fields, FP64) for every interior face between different orders, plus its
    120-byte face record."""
    q = desc.quads
    key = {}
    total = 0
    for e in range(q.shape[0]):
        for s_ in range(4):
            a, b = int(q[e, s_]), int(q[e, (s_ + 1) % 4])
            other = key.pop((b, a), None)
            if other is None:
                key[(a, b)] = e
            elif orders[other] != orders[e]:
                total += 2 * 24 * (orders[e] + 1 + orders[other] + 1) + 120
    return total


def cpu_baseline(desc, steps):
    """The oracle (the C restatement of the reference host path) timed on a bounded sample of the
    same workload: single-threaded (the reference's one MPI rank per core), then with OpenMP over
    the elements and faces of each stage on the cores this process may use (at most 16, one GPU's
    share of the host). `value` is the multi-threaded rate."""
    from oracle import oracle as O
    nproc = os.cpu_count() or 1
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    threads = max(1, min(16, avail))
    ot = O.Tables(N, 5)
    om = O.Mesh(ot, desc, N)
    rates = {}
    for nt, n_steps in ((1, steps), (threads, steps * 4)):
        used = O.set_threads(nt)
        om.initial_conditions(PULSES)
        t0 = time.perf_counter()
        om.run_steps(CFL, n_steps)
        rates[nt] = (3 * om.n * 3 * n_steps / (time.perf_counter() - t0), n_steps, used)
    O.set_threads(1)
    v1, s1, _ = rates[1]
    vt, st, used = rates[threads]
    return {"value": vt, "unit": "DOF*stage/s", "cores": used, "kind": "port",
            "single_thread": v1, "nproc": nproc, "cores_available": avail,
            "sample": f"the same {RES}x{RES} N={N} mesh, oracle/liboracle.so: {s1} RK3 steps on 1 thread "
                      f"({v1:.3g} DOF*stage/s), then {st} steps on {used} OpenMP threads over elements and faces"}


def batched_leg(args, nd, torch, stream):
    """Mixed orders 4..12 with mortar faces, FP64: direct launches with the default shapes, then
    autotuned shapes with one captured hipGraph per step (order groups as parallel branches)."""
    pulses = batched_workload()
    desc = nd.square_mesh(B_RES, jitter=0.0)
    desc.quad_N = batched_orders(desc, pulses)
    ndg = nd.NDG_t(12, 5)
    mesh = nd.Mesh2D_t(desc, 4, 12, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    mesh.set_stream(stream.cuda_stream)
    solver = nd.Solver2D_t(CFL, [0.0])
    dofs = mesh.n_dofs
    steps = args.steps

    def run(mode, tune, precision="fp64"):
        mesh.set_precision("fp64")
        mesh.initial_conditions(pulses)
        mesh.set_precision(precision)
        tuning = mesh.autotune(5) if tune else mesh.tuning()
        mesh.set_graph(mode)
        t = solver.run_steps(mesh, args.warmup, 0.0)
        mesh.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        solver.run_steps(mesh, steps, t)
        stop.record(stream)
        stop.synchronize()
        return start.elapsed_time(stop), tuning, mesh.get_solution()

    ms_direct, _, ref = run("off", False)
    ms_seq, _, got_seq = run("sequential", True)
    ms_graph, tuning, got = run("parallel", True)
    diff = max(float(np.abs(a - b).max()) for a, b in zip(got + got_seq, ref + ref))
    ms32, tuning32, got32 = run("parallel", True, "fp32c")
    mesh.set_precision("fp64")
    d32 = np.concatenate([a - b for a, b in zip(got32, got)])
    ref64 = np.concatenate(got)
    orders = desc.quad_N
    nf = orders.astype(np.int64) + 1
    bytes_launch = int((88 * nf * nf + 192 * nf + 80).sum()) + int(mortar_bytes(orders, desc))
    stage_s = ms_graph / 1e3 / (3 * steps)
    counts = {int(n): int((orders == n).sum()) for n in np.unique(orders)}
    return {"value": dofs * 3 * steps / (ms_graph / 1e3), "ms_per_step": ms_graph / steps,
            "value_direct_default_shapes": dofs * 3 * steps / (ms_direct / 1e3),
            "value_sequential_graph_tuned": dofs * 3 * steps / (ms_seq / 1e3),
            "workload": f"{B_RES}x{B_RES} walls, 8 seeded pulses, orders 4..12 by distance to the pulses, "
                        f"FP64, {steps} steps",
            "elements_per_order": counts, "dofs": dofs,
            "tuning": [{"N": n, "epb": e, "xcd_order": r} for n, e, r in tuning],
            "max_abs_diff_graph_vs_direct": diff,
            "fp32c": {"value": dofs * 3 * steps / (ms32 / 1e3), "ms_per_step": ms32 / steps,
                      "l2_rel_vs_fp64": float(np.linalg.norm(d32) / np.linalg.norm(ref64)),
                      "max_rel_vs_fp64": float(np.abs(d32).max() / np.abs(ref64).max()),
                      "tuning": [{"N": n, "epb": e, "xcd_order": r} for n, e, r in tuning32],
                      "note": "the same steps in the FP32 compensated form (mortar kernel in FP32 too), "
                              "tuned shapes, parallel graph branches"},
            "roofline": {"bound": "hbm", "achieved": bytes_launch / stage_s / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": bytes_launch / stage_s / 1e9 / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_stage": bytes_launch,
                         "note": "all launches of one stage (mortar kernel + one stage kernel per order)"}}


def adaptive_leg(nd, torch, stream, steps=1000, interval=20, graph="off"):
    """hp-adaptive run on the mixed-order path: 256 x 256 walls, N = 4 initially, orders up to 12,
    split level up to 1, the 8 seeded pulses, adaptation (estimate, split / raise, rebuild,
    transfer) every 20 steps, 1000 steps. Stepping throughput (tuned shapes; direct launches, or
    step graphs recaptured after each rebuild) and the adaptation time are reported separately."""
    pulses = batched_workload()
    desc = nd.square_mesh(B_RES, jitter=0.0)
    mesh = nd.Mesh2D_t(desc, 4, 12, 5, 1, 0, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(12, 5))
    mesh.set_stream(stream.cuda_stream)
    mesh.initial_conditions(pulses)
    mesh.set_graph(graph)
    solver = nd.Solver2D_t(CFL, [0.0])
    t, step_ms, adapt_s, tune_s, dof_stages = 0.0, 0.0, 0.0, 0.0, 0
    tuned, changed = set(), 0
    for block in range(steps // interval):
        if block:
            w0 = time.perf_counter()
            changed += mesh.adapt(12, 1)
            mesh.synchronize()
            adapt_s += time.perf_counter() - w0
        orders = set(int(n) for n in np.unique(mesh.orders_))
        if not orders <= tuned:
            w0 = time.perf_counter()
            mesh.autotune(3)
            tune_s += time.perf_counter() - w0
            tuned |= orders
        mesh.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        t = solver.run_steps(mesh, interval, t)
        stop.record(stream)
        stop.synchronize()
        step_ms += start.elapsed_time(stop)
        dof_stages += mesh.n_dofs * 3 * interval
        print(f"[adaptive] block {block}: K={mesh.n_elements_} dofs={mesh.n_dofs} t={t:.5f}", file=sys.stderr,
              flush=True)
    state = mesh.get_solution()
    if not all(np.isfinite(a).all() for a in state):
        raise SystemExit("bench.py: adaptive run blew up")
    counts = {int(n): int((mesh.orders_ == n).sum()) for n in np.unique(mesh.orders_)}
    out = {"value": dof_stages / (step_ms / 1e3), "unit": "DOF*stage/s",
           "workload": f"{B_RES}x{B_RES} walls, N0=4, N<=12, max split level 1, 8 seeded pulses, tol_min 1e-6, "
                       f"adapt every {interval} steps, {steps} steps, FP64",
           "stepping_s": step_ms / 1e3, "adaptation_s": adapt_s, "autotune_s": tune_s,
           "adapted_elements": changed, "final_elements": mesh.n_elements_, "final_dofs": mesh.n_dofs,
           "final_elements_per_order": counts, "t_final": t,
           "launches": graph,
           "note": "stepping time from HIP events around the 20-step blocks (with graphs, the capture after "
                   "each rebuild included); adaptation and tuning are host-synchronous wall times"}
    mesh.close()
    return out


# Equal-error workload: periodic [-1, 1]^2 with an exact solution (the Fourier series of the
# periodised pulses), the 8 seeded pulses of the batched workload, base mesh EE_RES^2, to t = EE_T.
EE_RES, EE_T, EE_INTERVAL = 128, 0.25, 20
EE_TARGETS = (1e-4, 1e-6, 1e-8, 1e-10)


def equal_error_leg(nd, torch, stream, res=EE_RES, t_end=EE_T, orders=(4, 6, 8, 10, 12),
                    tolerances=(1e-3, 1e-5, 1e-7, 1e-9, 1e-11, 1e-13), small_cfl=(0.35, 0.3, 0.25, 0.2, 0.125),
                    fp32c_tolerances=(1e-3, 1e-5), graph="off"):
    """The metric's comparison at equal L2 error. On the periodic multi-pulse workload, solved to
    t_end with CFL 0.5, every run's relative L2 error against the exact solution (Gauss quadrature
    of each element) and its time are measured for
      - plain: uniform order N (FP64, and the FP32 compensated form), for N in `orders`;
      - batched: the p-adaptive mixed-order path (N0 = 4, N <= 12, no splitting on the periodic
        square), adapt every 20 steps, after an iterative pre-conditioning of the initial
        condition (4 rounds of one step), for each estimator tolerance in `tolerances`; and the
        same path stepping in the FP32 compensated form (batched and mixed precision together)
        for the tolerances in `fp32c_tolerances` (its error floor is near 2e-7).
    At CFL 0.5 the RK3 time error (dt ~ 1/N^2) bounds every run near 3e-10 on this mesh, so the
    highest orders and the smallest tolerance also run at the CFL numbers in `small_cfl` (a grid
    fine enough that each form's cheapest run under 1e-10 is found near the CFL where its time
    error crosses it).
    For each target error, each form's cheapest run that reaches it is compared with the plain
    FP64 form's: time ratio, and the plain run's DOF*stage count divided by the form's time (the
    DOF*stage/s the plain path would need to reach the same error in the same time)."""
    from ndg_cuda_amd import exact
    pulses = batched_workload()
    desc = nd.square_mesh(res, x_periodic=True, y_periodic=True)
    ndg = nd.NDG_t(12, 5)
    solvers = {}

    def solver_for(cfl):
        if cfl not in solvers:
            solvers[cfl] = nd.Solver2D_t(cfl, [t_end])
        return solvers[cfl]

    def nodes_of(n):
        return ndg.nodes_[n], ndg.weights_[n]

    def error_of(mesh):
        X, Y, W = exact.mesh_nodes(desc, mesh.orders_, nodes_of)
        ex = exact.periodic_pulses(X, Y, pulses, t_end)
        return exact.l2_error(mesh.get_solution(), ex, W)

    def timed_solve(mesh, cfl, pre=None):
        """Pre-conditioning (if any), then the solve, between HIP events on the mesh stream; the
        adaptations inside are host-synchronous and counted in the elapsed time."""
        solver = solver_for(cfl)
        mesh.synchronize()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        wall0 = time.perf_counter()
        start.record(stream)
        if pre:
            pre()
        solver.solve(None, mesh)
        stop.record(stream)
        stop.synchronize()
        return start.elapsed_time(stop) / 1e3, time.perf_counter() - wall0

    runs = []

    def plain(n, cfl, precisions):
        mesh = nd.Mesh2D_t(desc, n, n, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
        mesh.set_stream(stream.cuda_stream)
        mesh.set_graph(graph)
        for prec in precisions:
            mesh.set_precision("fp64")
            mesh.initial_conditions(pulses, period=(2.0, 2.0))
            mesh.set_precision(prec)
            s, wall = timed_solve(mesh, cfl)
            mesh.set_precision("fp64")
            steps = solver_for(cfl).n_steps_
            runs.append({"form": "plain_" + prec, "N": n, "cfl": cfl, "seconds": s, "wall_s": wall, "steps": steps,
                         "dof_stages": mesh.n_dofs * 3 * steps, "l2_error": error_of(mesh)})
            print(f"[equal-error] {runs[-1]}", file=sys.stderr, flush=True)
        mesh.close()

    def batched(tol, cfl, prec="fp64"):
        mesh = nd.Mesh2D_t(desc, 4, 12, 5, 0, EE_INTERVAL, 0, tol, 1e-14, 0.0, ndg)
        mesh.set_stream(stream.cuda_stream)
        mesh.set_graph(graph)
        mesh.initial_conditions(pulses, period=(2.0, 2.0))
        solver = solver_for(cfl)

        def pre():
            solver.pre_condition_iterative(ndg, mesh, 4, 1)
            mesh.set_precision(prec)

        s, wall = timed_solve(mesh, cfl, pre=pre)
        mesh.set_precision("fp64")
        counts = {int(k): int((mesh.orders_ == k).sum()) for k in np.unique(mesh.orders_)}
        runs.append({"form": "batched_" + prec, "tolerance_min": tol, "cfl": cfl, "seconds": s, "wall_s": wall,
                     "steps": solver.n_steps_, "final_dofs": mesh.n_dofs, "final_elements_per_order": counts,
                     "l2_error": error_of(mesh)})
        print(f"[equal-error] {runs[-1]}", file=sys.stderr, flush=True)
        mesh.close()

    for n in orders:
        plain(n, CFL, ("fp64", "fp32c"))
    for tol in tolerances:
        batched(tol, CFL)
    for tol in fp32c_tolerances:
        batched(tol, CFL, "fp32c")
    for cfl in small_cfl:
        for n in orders[-2:]:
            plain(n, cfl, ("fp64",))
        batched(tolerances[-1], cfl)
    table = []
    for target in EE_TARGETS:
        row = {"target_l2": target}
        best = {}
        for r in runs:
            if r["l2_error"] <= target and (r["form"] not in best or r["seconds"] < best[r["form"]]["seconds"]):
                best[r["form"]] = r
        plain = best.get("plain_fp64")
        for form, r in sorted(best.items()):
            entry = {"seconds": r["seconds"], "l2_error": r["l2_error"]}
            entry.update({k: r[k] for k in ("N", "tolerance_min", "cfl") if k in r})
            if plain is not None:
                entry["speedup_vs_plain_fp64"] = plain["seconds"] / r["seconds"]
                entry["equivalent_dof_stage_per_s"] = plain["dof_stages"] / r["seconds"]
            row[form] = entry
        table.append(row)
    return {"workload": f"{res}x{res} periodic, 8 seeded pulses (widths 0.05-0.2) periodised with period 2, "
                        f"CFL {CFL} (and {', '.join(str(c) for c in small_cfl)} for the highest orders and the "
                        f"smallest tolerance), to t = {t_end}; "
                        "error = relative L2 (Gauss quadrature) against the exact Fourier-series solution",
            "runs": runs, "at_equal_error": table,
            "note": "seconds = HIP events around pre-conditioning + solve on the mesh stream (host adaptation "
                    "included); speedup = plain FP64 time / form time, each form's cheapest run reaching the target"}


def profile_batched(steps, graph):
    """The mixed-order workload alone, for kernel counters: IC, then `steps` RK3 steps (direct
    launches with the default shapes, or tuned shapes in captured step graphs); prints the
    algorithmic bytes of each order group's stage launch and of the mortar kernel."""
    import ndg_cuda_amd as nd
    pulses = batched_workload()
    desc = nd.square_mesh(B_RES, jitter=0.0)
    desc.quad_N = batched_orders(desc, pulses)
    mesh = nd.Mesh2D_t(desc, 4, 12, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(12, 5))
    mesh.initial_conditions(pulses)
    if graph != "off":
        mesh.autotune(3)
        mesh.set_graph(graph)
    nd.Solver2D_t(CFL, [0.0]).run_steps(mesh, steps, 0.0)
    mesh.synchronize()
    orders = desc.quad_N
    groups = {int(n): {"elements": int((orders == n).sum()),
                       "bytes_per_launch": int((orders == n).sum()) * algorithmic_bytes_per_element(int(n))}
              for n in np.unique(orders)}
    print(json.dumps({"groups": groups, "mortar_bytes_per_launch": int(mortar_bytes(orders, desc)),
                      "tuning": [[int(v) for v in t] for t in mesh.tuning()], "steps": steps, "graph": graph}))


def timed_steps(solver, mesh, stream, torch, steps, t=0.0):
    """Runs `steps` RK3 steps on the mesh stream between two HIP events; returns ms."""
    mesh.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    solver.run_steps(mesh, steps, t)
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop)


def fp32c_periodic_16x16(nd):
    """16 x 16 periodic, N = 6, one pulse, CFL 0.5 to t = 0.5 (289 steps): the FP32 compensated
    form's relative l2 distance to the FP64 run (which matches the reference to 1e-12)."""
    desc = nd.square_mesh(16, x_periodic=True, y_periodic=True)
    mesh = nd.Mesh2D_t(desc, 6, 6, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(6, 5))
    states = {}
    for prec in ("fp64", "fp32c"):
        mesh.set_precision("fp64")
        mesh.initial_conditions([(0.1, -0.2, 0.2, 1.0)])
        mesh.set_precision(prec)
        nd.Solver2D_t(CFL, [0.5]).solve(None, mesh)
        states[prec] = np.concatenate(mesh.get_solution())
    mesh.close()
    d = states["fp32c"] - states["fp64"]
    return {"workload": "16x16 periodic, N=6, 1 pulse, t=0.5 (289 steps)",
            "l2_rel_vs_fp64": float(np.linalg.norm(d) / np.linalg.norm(states["fp64"])),
            "max_rel_vs_fp64": float(np.abs(d).max() / np.abs(states["fp64"]).max())}


def precision_leg(nd, torch, stream):
    """1024 x 1024 periodic, N = 7, 4 seeded pulses, 500 steps: FP64 and the FP32 compensated
    form on the same initial state, with the FP32 form's error against the FP64 run."""
    res, n, steps = 1024, 7, 500
    rng = np.random.default_rng(4)
    pulses = np.column_stack([rng.uniform(-0.7, 0.7, 4), rng.uniform(-0.7, 0.7, 4), rng.uniform(0.1, 0.2, 4),
                              rng.uniform(0.2, 1.0, 4)])
    desc = nd.square_mesh(res, x_periodic=True, y_periodic=True)
    mesh = nd.Mesh2D_t(desc, n, n, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(n, 5))
    mesh.set_stream(stream.cuda_stream)
    solver = nd.Solver2D_t(CFL, [0.0])
    out = {"workload": f"{res}x{res} periodic, N={n}, 4 seeded pulses, {steps} steps", "dofs": mesh.n_dofs}
    states = {}
    for prec in ("fp64", "fp32c"):
        mesh.set_precision("fp64")
        mesh.initial_conditions(pulses)
        mesh.set_precision(prec)
        ms = timed_steps(solver, mesh, stream, torch, steps)
        states[prec] = np.concatenate(mesh.get_solution())
        nf = n + 1
        b = algorithmic_bytes_per_element(n) if prec == "fp64" else algorithmic_bytes_fp32c(n)
        b *= mesh.n_elements_
        out[prec] = {"value": mesh.n_dofs * 3 * steps / (ms / 1e3), "ms_per_step": ms / steps,
                     "roofline_frac": b / (ms / 1e3 / (3 * steps)) / 1e9 / HBM_PEAK_GBS}
    d = states["fp32c"] - states["fp64"]
    out["fp32c"]["l2_rel_vs_fp64"] = float(np.linalg.norm(d) / np.linalg.norm(states["fp64"]))
    out["fp32c"]["max_rel_vs_fp64"] = float(np.abs(d).max() / np.abs(states["fp64"]).max())
    mesh.close()
    return out


def adjoint_leg(nd, torch, stream, steps=2000):
    """128 x 128 periodic, N = 8, 2000 steps, checkpoints every 50, 32 seeded receivers, data from a
    c = 1.05 run of two pulses: gradient time against one forward solve, and one central finite
    difference (eps = 1e-6) along a seeded direction. (Fewer steps only for profiler runs.)"""
    from ndg_cuda_amd import adjoint as A
    res, n, interval = 128, 8, 50
    desc = nd.square_mesh(res, x_periodic=True, y_periodic=True)
    mesh = nd.Mesh2D_t(desc, n, n, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(n, 5))
    mesh.set_stream(stream.cuda_stream)
    rec = A.Receivers.seeded(mesh.n_elements_, 32, 2024)
    dt = 0.5 * (2.0 / res) / (n * n)
    mesh.initial_conditions([(0.1, -0.2, 0.2, 1.0), (-0.4, 0.3, 0.2, 0.5)])
    A.set_wave_speed(mesh, 1.05)
    data = A.record_receivers(mesh, rec, dt, steps)
    A.set_wave_speed(mesh, 1.0)
    mesh.initial_conditions([(0.1, -0.2, 0.2, 1.0)])
    q0 = [np.array(a) for a in mesh.get_solution()]
    t0 = time.perf_counter()
    A.record_receivers(mesh, rec, dt, steps)
    t_fwd = time.perf_counter() - t0

    def gradient():
        """One gradient call from q0, timed on its second run: the first sizes the mesh's
        checkpoint and stage-store blocks (kept for later calls), as the forward solve timed
        above is the second recorded run."""
        for _ in range(2):
            mesh.set_solution(*q0)
            t0 = time.perf_counter()
            out = A.adjoint_gradient(mesh, rec, dt, steps, interval, data)
            t = time.perf_counter() - t0
        return t, out

    t_adj, (J, gp, gu, gv, gc) = gradient()
    seg_bytes = 3 * interval * 3 * mesh.n_dofs_per_field_ * 8
    n_seg = (steps + interval - 1) // interval
    free, _ = torch.cuda.mem_get_info()

    def budgeted(n_kept):
        """The same gradient with the stage states of the last n_kept segments kept from the forward
        pass (their recompute skipped); returns (seconds, identical to the checkpointed one)."""
        A.set_store_budget(mesh, (n_kept - 1) * seg_bytes)
        t2, (J2, gp2, gu2, gv2, gc2) = gradient()
        A.set_store_budget(mesh, 0)
        same = J2 == J and gc2 == gc and all(np.array_equal(a, b) for a, b in zip((gp2, gu2, gv2), (gp, gu, gv)))
        return t2, bool(same)

    # every segment kept (no recompute), when the stage states fit in the free device memory (the
    # blocks stay with the mesh for the next call)
    kept = {"fits": False, "bytes": n_seg * seg_bytes}
    if n_seg * seg_bytes + (16 << 30) < free:
        t_kept, same = budgeted(n_seg)
        kept.update({"fits": True, "gradient_s": t_kept, "gradient_over_forward": t_kept / t_fwd,
                     "identical_to_checkpointed": same})
    # checkpoints every 50 steps and recompute, with the last 16 of the 40 segments kept (77 GB, a
    # quarter of the HBM; reuses the blocks of the call above)
    part = {"segments_kept": min(16, n_seg), "bytes": min(16, n_seg) * seg_bytes, "fits": False}
    if part["bytes"] + (16 << 30) < free:
        t_part, same = budgeted(part["segments_kept"])
        part.update({"fits": True, "gradient_s": t_part, "gradient_over_forward": t_part / t_fwd,
                     "identical_to_checkpointed": same})
    # dJ/dc from the stage right-hand sides recorded in the forward pass: no checkpoints, no recompute.
    A.set_c_derivative(mesh, "stage_records")
    t_rec, (J3, gp3, gu3, gv3, gc3) = gradient()
    A.set_c_derivative(mesh, "recompute")
    stage_records = {"gradient_s": t_rec, "gradient_over_forward": t_rec / t_fwd, "dJ_dc": gc3,
                     "dJ_dc_rel_diff_vs_recompute": abs(gc3 - gc) / abs(gc),
                     "initial_field_gradients_identical": bool(
                         J3 == J and all(np.array_equal(a, b) for a, b in zip((gp3, gu3, gv3), (gp, gu, gv))))}
    rng = np.random.default_rng(8)
    d = [rng.normal(size=q0[0].size) for _ in range(3)]
    eps = 1e-6

    def misfit(sign):
        mesh.set_solution(*(a + sign * eps * b for a, b in zip(q0, d)))
        A.set_wave_speed(mesh, 1.0 + sign * eps)
        p = A.record_receivers(mesh, rec, dt, steps)
        return 0.5 * float(((p - data) ** 2).sum())

    fd = (misfit(1) - misfit(-1)) / (2 * eps)
    ad = float(gp @ d[0] + gu @ d[1] + gv @ d[2]) + gc
    ad3 = float(gp3 @ d[0] + gu3 @ d[1] + gv3 @ d[2]) + gc3
    stage_records["fd_rel_diff"] = abs(ad3 - fd) / abs(fd)
    mesh.close()
    ratios = {"checkpoint_recompute": t_adj / t_fwd, "stage_records": t_rec / t_fwd}
    if kept["fits"]:
        ratios["all_segments_kept"] = kept["gradient_over_forward"]
    if part["fits"]:
        ratios["sixteen_segments_kept"] = part["gradient_over_forward"]
    best = min(ratios, key=ratios.get)
    return {"workload": f"{res}x{res} periodic, N={n}, {steps} steps, checkpoints every {interval}, 32 receivers",
            "J": J, "dJ_dc": gc, "forward_s": t_fwd, "gradient_s": t_adj, "gradient_over_forward": t_adj / t_fwd,
            "gradient_over_forward_is": "checkpoints every 50 steps, every segment but the last recomputed",
            "gradient_over_forward_best": ratios[best], "best_method": best,
            "all_segments_kept": kept, "sixteen_segments_kept": part,
            "c_derivative_from_stage_records": stage_records,
            "fd_rel_diff": abs(ad - fd) / abs(fd),
            "note": "wall times of the library calls (host-synchronous), forward = one recorded 2000-step run "
                    "(the second), each gradient timed on its second call (the first sizes the mesh's "
                    "checkpoint and stage-store blocks)"}


def copy_rate(torch, stream, gib=2, reps=10):
    """Device-to-device copy rate (bytes read + written per second) of a plain torch copy of
    `gib` GiB, timed with HIP events on the bench stream: the practical HBM rate on this GPU."""
    n = (gib << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    with torch.cuda.stream(stream):
        b.copy_(a)
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(reps):
            b.copy_(a)
        stop.record(stream)
    stop.synchronize()
    rate = 2 * n * 8 * reps / (start.elapsed_time(stop) / 1e3)
    del a, b
    torch.cuda.empty_cache()
    return rate


def separate_forms(args, legs, nd, torch, stream):
    """The form legs that run on workloads of their own (the FP32 leg reuses the main mesh)."""
    forms = {}
    if "batched" in legs:
        forms["batched"] = batched_leg(args, nd, torch, stream)
    if "precision" in legs:
        forms["precision_1024"] = precision_leg(nd, torch, stream)
    if "adjoint" in legs:
        forms["adjoint_128"] = adjoint_leg(nd, torch, stream, args.adjoint_steps)
    if "equal_error" in legs:
        forms["equal_error"] = equal_error_leg(nd, torch, stream)
    if "adaptive" in legs:
        forms["adaptive_256"] = adaptive_leg(nd, torch, stream)
        forms["adaptive_256"]["value_parallel_graphs"] = adaptive_leg(nd, torch, stream, graph="parallel")["value"]
    return forms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="RK3 steps of the single-thread CPU baseline (4x as many multi-threaded)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-only", action="store_true", help="skip every form leg (same as --forms none)")
    ap.add_argument("--forms", default="batched,adaptive,precision,adjoint,fp32c,equal_error",
                    help="comma-separated form legs to run after the main line (or 'none')")
    ap.add_argument("--forms-only", action="store_true",
                    help="profiler runs: skip the main line and print only the selected form legs")
    ap.add_argument("--adjoint-steps", type=int, default=2000,
                    help="time steps of the adjoint leg (2000 is the workload; fewer for profiler runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (reference for HBM counters)")
    ap.add_argument("--profile-batched", choices=("off", "sequential", "parallel"),
                    help="profiler runs: only the mixed-order workload, --steps steps with direct launches "
                         "and the default shapes (off), or tuned shapes in captured step graphs")
    args = ap.parse_args()
    if args.profile_batched:
        profile_batched(args.steps, args.profile_batched)
        return
    if args.copy_calibration:
        import torch
        n = (1 << 30) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); the DG stage "
              "runs on a single MI355X and multi-GPU sharding is not part of it", file=sys.stderr)
        sys.exit(2)

    import torch
    import ndg_cuda_amd as nd

    if nd.device_count() < 1 or not torch.cuda.is_available():
        raise SystemExit("bench.py: no GPU visible")
    stream = torch.cuda.Stream()
    legs = set() if args.fp64_only else set(args.forms.split(","))
    if args.forms_only:
        print(json.dumps({"forms": separate_forms(args, legs, nd, torch, stream)}))
        return
    desc = nd.square_mesh(RES, jitter=JITTER, seed=SEED)
    ndg = nd.NDG_t(N, 5)
    mesh = nd.Mesh2D_t(desc, N, N, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    mesh.set_stream(stream.cuda_stream)
    solver = nd.Solver2D_t(CFL, [0.0])

    # The library's own launch-shape tuner (per precision) and one captured hipGraph per RK step.
    tuned = {}
    mesh.set_graph("sequential")

    def timed_run(precision):
        """IC, W warm-up steps, K timed steps; returns (elapsed ms, wall s, final state)."""
        mesh.set_precision("fp64")
        mesh.initial_conditions(PULSES)
        mesh.set_precision(precision)
        if precision not in tuned:
            tuned[precision] = mesh.autotune(5)[0]
        dt = solver.get_delta_t(mesh)
        t = 0.0
        for _ in range(args.warmup):
            solver.step(mesh, t, dt)
            t += dt
        mesh.synchronize()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        wall0 = time.perf_counter()
        start.record(stream)
        for _ in range(args.steps):
            solver.step(mesh, t, dt)
            t += dt
        stop.record(stream)
        stop.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        state = mesh.get_solution()
        if not all(a.size and float(abs(a).max()) < 1e3 for a in state):
            raise SystemExit(f"bench.py: {precision} solution blew up")
        return start.elapsed_time(stop), wall, state

    dofs = mesh.n_dofs
    stages = 3 * args.steps
    ms, wall, ref_state = timed_run("fp64")
    value = dofs * stages / (ms / 1e3)
    launch_s = ms / 1e3 / stages
    bytes_per_launch = algorithmic_bytes_per_element(N) * mesh.n_elements_
    achieved = bytes_per_launch / launch_s / 1e9
    copy_gbs = copy_rate(torch, stream) / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    forms = separate_forms(args, legs, nd, torch, stream)
    traffic_c32 = None
    if os.path.exists(TRAFFIC_FILE_C32):
        with open(TRAFFIC_FILE_C32) as f:
            traffic_c32 = json.load(f).get("hbm_bytes_per_launch")
    if "fp32c" in legs:
        ms32, _, st32 = timed_run("fp32c")
        diff = np.concatenate([a - b for a, b in zip(st32, ref_state)])
        refv = np.concatenate(ref_state)
        b32 = algorithmic_bytes_fp32c(N) * mesh.n_elements_
        forms["fp32c"] = {
            "value": dofs * stages / (ms32 / 1e3), "ms_per_step": ms32 / args.steps,
            "l2_rel_vs_fp64": float(np.linalg.norm(diff) / np.linalg.norm(refv)),
            "max_rel_vs_fp64": float(np.abs(diff).max() / np.abs(refv).max()),
            "roofline": {"bound": "hbm", "achieved": b32 / (ms32 / 1e3 / stages) / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": b32 / (ms32 / 1e3 / stages) / 1e9 / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_launch": b32, "kernel": "ndg::stage_c32_kernel<8>",
                         "traffic": traffic_c32},
            "note": "FP32 right-hand side and RK register, state as a two-sum-compensated float pair; "
                    "error is against the FP64 run of the same steps",
            "periodic_16x16_N6": fp32c_periodic_16x16(nd)}
    out = {
        "metric": "DOF\u00b7stage updates/s: FP64 plain vs batched/mixed-prec at equal L2 error",
        "value": value,
        "value_is": "the FP64 plain stage; the batched and mixed-precision forms, with their l2 "
                    "distance to FP64, are under forms",
        "unit": "DOF*stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{RES}x{RES} quads jittered +-{int(JITTER * 100)}%, Hilbert order, N={N}, "
                               f"walls, 2 Gaussian pulses, FP64, one RK3 step = 3 fused stage launches",
                   "elements": mesh.n_elements_, "N": N, "dofs": dofs, "stage_launches": stages,
                   "parallelism": "single GPU",
                   "launch": {"elements_per_workgroup": tuned["fp64"][1], "xcd_order": tuned["fp64"][2],
                              "step_graph": "sequential", "shape_from": "ndg_mesh_autotune"}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "ndg::stage_kernel<8,false>", "avg_launch_ms": launch_s * 1e3,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "copy_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs,
                     "hbm_gbs": traffic / launch_s / 1e9 if traffic else None,
                     "hbm_frac": traffic / launch_s / 1e9 / HBM_PEAK_GBS if traffic else None},
        "forms": forms,
        "wall_s_timed_region": wall,
    }
    if not args.no_cpu and args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(desc, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
