#!/usr/bin/env python3
"""Benchmark: DoF-updates/s of the ADER-DG cell update, 3D Euler order 5.

Workload (one GPU): 3D compressible Euler ADER-DG, order N=5 (6^3 nodes per
cell), 5 variables, 81^3 periodic cells on the unit cube, smooth density wave
plus 8 seeded random-phase Fourier modes (amplitude 1e-2, |k_i| <= 4) in rho
and u, global time stepping (CFL 0.9), K time steps.

Two forms run in the same process on the same initial state:
  * FP64  -- every kernel in FP64, adaptive Picard (tol 1e-7), parity-tested
             against the CPU oracle at <= 1e-12;
  * FP32  -- space-time predictor in FP32, fixed Picard count (the FP32
             residual cannot reach the absolute 1e-7 tolerance), Riemann,
             corrector, solution update and CFL reduction in FP64.
`value` is the FP32 form's DoF-updates/s; the FP64 form, the per-variable
relative L2 error of FP32 vs FP64 and the speed-up are reported beside it.
With --slabs S the FP32 form's timed steps run slab-pipelined (S slabs of z
planes: the corrector of a slab on a second stream beside the predictors of
the later slabs, bitwise the one-launch step's results); its per-kernel times
and the roofline then come from the same steps re-run with one launch per
phase.  Default 0: one launch per phase (the pipelined step measured within
0-2 % of it).

`--workload batched` measures the batched uneven workload: a static
two-level 3D Euler mesh of order 4 (27^3 coarse cells, those within 0.25 of
the centre refined 3x, one global dt) plus 64 independent 2D Euler meshes of
order 4 with seeded sizes between 9^2 and 81^2 cells (one dt each), stepped
together by one captured hipGraph per step after the autotuner has chosen the
launch geometry of every member.

`--workload acoustic` instead measures the linear acoustic workload (3D,
order 4, 54^3 cells, per-cell wave speed with 8 Gaussian bumps, Ricker
source f0 = 5, 64 receivers, 300 steps): forward DoF-updates/s in FP64 and
the wall time of the misfit gradient (discrete adjoint, checkpoints every 20
steps).

A DoF-update is one (node, variable) advanced by one time step:
cells * (N+1)^3 * nVar * steps / wall time.  The timed region starts with the
state resident in HBM and brackets exactly K steps with device syncs.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "exahype-engine_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
HBM_PEAK_GBS = 8000.0
# HBM bytes per predictor launch at 81^3, order 5 (n, order, precision, CERK guess), measured with
# rocprofv3 FETCH_SIZE (x2, the factor a 1 GiB copy calibrates) + WRITE_SIZE (x1):
# profiles/euler3d_n5_81cubed_kernels.md.
MEASURED_PREDICTOR_TRAFFIC = {(81, 5, "fp32", 1): 14.68e9, (81, 5, "fp32", 0): 17.85e9,
                              (81, 5, "fp64", 1): 18.46e9, (81, 5, "fp64", 0): 18.73e9}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def predictor_flops_per_cell(dim, B, V, ND, iters, cerk=False):
    """Algorithmic contraction flops of one predictor+volume-integral call.

    Per Picard iteration: D stiffness contractions 2*V*B^(D+2) each, the iK1
    time solve 2*V*B^(D+2), rhs0 + residual ~5*V*B^(D+1).  Once per call: time
    averages of lQi and the D fluxes 2*B^(D+1)*(ND + D*V), extrapolation to
    2D faces 2*2D*B^(D-1)*B*(ND + V), volume integral D*2*V*B^(D+1).
    Flux evaluations are not counted (the reference's counting convention,
    Miscellaneous/KernelsPerfMeasurementsTests/report/report.tex:395-408).
    CERK initial guess: two flux divergences 2 * D*2*V*B^(D+1) (derivative
    contractions at every node) and the guess polynomial ~6*V*B^(D+1)."""
    per_iter = dim * 2 * V * B ** (dim + 2) + 2 * V * B ** (dim + 2) + 5 * V * B ** (dim + 1)
    once = 2 * B ** (dim + 1) * (ND + dim * V) + 2 * 2 * dim * B ** (dim - 1) * B * (ND + V) + dim * 2 * V * B ** (dim + 1)
    if cerk:
        once += 2 * dim * 2 * V * B ** (dim + 1) + 6 * V * B ** (dim + 1)
    return per_iter * iters + once


def equal_slice_savings_per_cell(dim, B, V, iters):
    """Contraction flops the node-column kernel does not execute with the trivial
    guess: the first iterate is constant in time, so its D stiffness contractions
    run for one time slice instead of B (cells that iterate at least once)."""
    return np.where(np.asarray(iters) > 0, dim * 2 * V * B ** (dim + 1) * (B - 1), 0)


def kind_is_node_column(precision, variant, E):
    """Whether the mesh runs the node-column kernel for 3D order 5: FP64 with the
    default variant (0) or variant 1 or 4, and any precision with variant 1; the
    line / tile kernels otherwise."""
    return (precision == E.PREC_FP64 and variant in (0, 1, 4)) or variant == 1


def executed_iterations(counts, dim, max_iter, adaptive):
    counts = np.asarray(counts)
    if not adaptive:
        return np.full(counts.shape, max_iter)
    if dim == 3:  # 0-based index of the converged iteration, max_iter if not converged
        return np.where(counts < max_iter, counts + 1, max_iter)
    return np.where(counts == max_iter + 2, counts, max_iter)  # 2D convention loses the index


def run_form(E, P, n, order, steps, warmup, precision, picard, seed, modes, variant=0, guess=0, slabs=0):
    dim, V = 3, 5
    ctx = E.Context(dim, order, V, 0, precision=precision, picard_max=picard)
    ctx.set_predictor_variant(variant)
    ctx.set_initial_guess(guess)
    mesh = E.Mesh(ctx, n, 1.0 / n)
    if slabs:
        mesh.set_overlap(slabs)  # slab-pipelined step (runs with timing events use the one-launch step)
    prm = P.density_wave_params(dim, seed, modes)
    # warm-up steps on the same state, then restart from the initial condition
    mesh.adjust_solution(prm)
    mesh.init_dt(0.0)
    if warmup:
        mesh.run(warmup)
        mesh.sync()
    mesh.adjust_solution(prm)
    mesh.init_dt(0.0)
    if not slabs:
        mesh.set_timing(True)  # HIP events around each kernel on the launch stream, no syncs
    mesh.sync()
    t0 = time.perf_counter()
    mesh.run(steps)
    mesh.sync()
    wall = time.perf_counter() - t0
    pic = mesh.picard()
    recomputed = mesh.picard_guard_recomputed()
    t_end, dt_next = mesh.time()
    sol = mesh.get_solution()
    one_launch = None
    if slabs:
        # per-kernel times: the same steps from the same state in the one-launch step with HIP
        # events around each kernel (the kernels' own rates, not shared with the other stream)
        mesh.adjust_solution(prm)
        mesh.init_dt(0.0)
        mesh.set_timing(True)
        mesh.sync()
        t1 = time.perf_counter()
        mesh.run(steps)
        mesh.sync()
        one_launch = dict(ms_per_step=(time.perf_counter() - t1) * 1e3 / steps,
                          bitwise_equal=bool(np.array_equal(mesh.get_solution(), sol)))
    ph = mesh.phase_times()
    mesh.close()
    ctx.close()
    B = order + 1
    cells = n ** dim
    adaptive = picard <= 0
    max_iter = 2 * (order + 1) if adaptive else picard
    iters = executed_iterations(pic, dim, max_iter, adaptive)
    # algorithmic predictor flops of one launch at the Picard counts of the last
    # step (the algorithmic contraction model of predictor_flops_per_cell); the CERK guess's two flux
    # divergences are reported apart
    flops = float(np.sum(predictor_flops_per_cell(dim, B, V, V, iters)))
    cerk_flops = float(np.sum(predictor_flops_per_cell(dim, B, V, V, iters, cerk=True))) - flops if guess else 0.0
    # contraction flops the kernel executes: the node-column kernel (the FP64 form, FP32
    # variant 1) forms the trivial guess's first iteration from one time slice
    node_column = kind_is_node_column(precision, variant, E)
    executed = flops - (float(np.sum(equal_slice_savings_per_cell(dim, B, V, iters)))
                        if node_column and not guess else 0.0)
    pred_ms = ph["predictor_ms"] / steps
    kind = {E.PREC_FP64: "fp64", E.PREC_FP32: "fp32", E.PREC_BF16: "bf16"}[precision]
    achieved = flops / (pred_ms * 1e-3) / 1e12
    dofs = cells * B ** dim * V * steps
    if kind == "fp64":
        bound = "fp64-valu"
    elif kind == "bf16":
        bound = "bf16-mfma"
    elif picard == 1 and variant == 0 and dim == 3 and B == 6:
        bound = "fp32-valu"  # one-iteration line kernel: vector ALU only
    else:
        bound = "fp32-valu+mfma"
    return dict(
        kind=kind, wall_s=wall, ms_per_step=wall * 1e3 / steps, value=dofs / wall,
        predictor_ms=pred_ms, riemann_ms=ph["riemann_ms"] / steps, corrector_ms=ph["corrector_ms"] / steps,
        picard_hist={int(k): int(v) for k, v in zip(*np.unique(iters, return_counts=True))},
        guard_recomputed_cells=int(recomputed),
        picard_max_executed=int(iters.max()), predictor_flops=flops, cerk_flops=cerk_flops,
        roofline=dict(bound=bound, kernel="predictor", achieved=achieved, peak=PEAK_TFLOPS[kind],
                      unit="TFLOP/s", frac=achieved / PEAK_TFLOPS[kind],
                      traffic=MEASURED_PREDICTOR_TRAFFIC.get((n, order, kind, int(bool(guess)))),
                      flops_per_launch=flops, cerk_flops_per_launch=cerk_flops,
                      executed_flops_per_launch=executed,
                      achieved_executed=executed / (pred_ms * 1e-3) / 1e12,
                      kernel_ms=pred_ms,
                      effective_with_cerk=(flops + cerk_flops) / (pred_ms * 1e-3) / 1e12),
        t_end=t_end, dt_next=dt_next, solution=sol, slabs=slabs, one_launch_step=one_launch)


def cpu_threads():
    """(threads used, CPUs visible): the threads are the host's CPU share of one
    GPU (16 on the GPU box), capped by what this process may run on."""
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count() or 1
    return max(1, min(16, visible)), visible


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(O, P, order, n_cpu, steps_cpu, threads, visible, n_one=18, steps_one=1):
    """The oracle on the host: `threads` OpenMP threads on n_cpu^3 cells, and one
    thread on n_one^3 cells (the per-core figure)."""
    orc = O.Oracle(3, order)
    luh = P.density_wave_3d(n_cpu, order)
    t0 = time.perf_counter()
    orc.run_uniform(luh, n_cpu, 1.0 / n_cpu, steps_cpu, 0.0, threads=threads)
    t = time.perf_counter() - t0
    dofs = n_cpu ** 3 * (order + 1) ** 3 * 5 * steps_cpu
    luh1 = P.density_wave_3d(n_one, order)
    t1 = time.perf_counter()
    orc.run_uniform(luh1, n_one, 1.0 / n_one, steps_one, 0.0, threads=1)
    t1 = time.perf_counter() - t1
    dofs1 = n_one ** 3 * (order + 1) ** 3 * 5 * steps_one
    return dict(value=dofs / t, unit="DoF-updates/s", cores=threads, kind="port",
                value_1core=dofs1 / t1, cpu_model=cpu_model(), cpus_visible=visible,
                sample=f"oracle (FP64 C restatement of the reference generic kernels, OpenMP over cells) on "
                       f"{n_cpu}^3 periodic cells, order {order}, {steps_cpu} global steps of the same workload "
                       f"({t:.2f} s wall, {threads} threads = the host's CPU share of one GPU, {visible} CPUs "
                       f"visible); 1-core figure: {n_one}^3 cells, {steps_one} step(s), one thread ({t1:.2f} s)")


def acoustic_bytes_per_cell(B):
    """Algorithmic HBM bytes per cell and step of the two acoustic kernels.
    predictor: read u (4 doubles per node) + material (2), write u and the six
    extrapolated faces (4 doubles per face node); corrector: read u, its own
    faces and the six neighbour faces (each face is read by both of its
    cells), material, write u."""
    N, F = B ** 3, B ** 2
    pred = 8 * (4 * N + 2 + 4 * N + 6 * F * 4)
    corr = 8 * (4 * N + 6 * F * 4 + 6 * F * 4 + 2 + 4 * N)
    return pred, corr


def acoustic_cpu_baseline(O, P, order, n_cpu, steps_cpu, threads):
    c = P.acoustic_wave_speed(n_cpu, seed=0)
    luh = np.ascontiguousarray(P.acoustic_state(n_cpu, order, c))
    h = 1.0 / n_cpu
    dt = 0.9 * O.PNPM[order] * h / (3 * c.max())
    o = O.LinearOracle(3, order)
    src = (np.array(P.ACOUSTIC_SOURCE), P.ACOUSTIC_AMP, P.ACOUSTIC_F0, P.ACOUSTIC_T0)
    t0 = time.perf_counter()
    o.run(luh, n_cpu, h, dt, steps_cpu, src, P.acoustic_receivers(seed=0), threads=threads)
    t = time.perf_counter() - t0
    dofs = n_cpu ** 3 * (order + 1) ** 3 * 4 * steps_cpu
    return dict(value=dofs / t, unit="DoF-updates/s", cores=threads, kind="port",
                sample=f"oracle (FP64 C restatement of the reference generic linear kernels, OpenMP over cells) on "
                       f"{n_cpu}^3 periodic cells, order {order}, {steps_cpu} steps of the same workload "
                       f"({t:.2f} s wall, {threads} threads)")


def bench_acoustic(args, E, P):
    n, order, steps, every = args.n, args.order, args.steps, 20
    B = order + 1
    ctx = E.Context(3, order, nvar=4, npar=2, pde=E.PDE_ACOUSTIC)
    c_true = P.acoustic_wave_speed(n, seed=args.seed, scale=1.0)
    c0 = P.acoustic_wave_speed(n, seed=args.seed, scale=0.5)
    rec = P.acoustic_receivers(seed=args.seed, nrec=64)

    def make(c, dt=0.0):
        ac = E.Acoustic(ctx, n, 1.0 / n)
        ac.set_predictor(args.acoustic_predictor)
        ac.set_material(rho=np.ones(n ** 3), c=c)
        ac.set_dt(dt)
        ac.set_source(np.array(P.ACOUSTIC_SOURCE), P.ACOUSTIC_AMP, P.ACOUSTIC_F0, P.ACOUSTIC_T0)
        ac.set_receivers(rec)
        return ac

    truth = make(c_true)
    dt = truth.time()[2]
    if args.warmup:
        truth.run(args.warmup)
        truth.set_step(0)
        truth.set_material(c=c_true)
        truth.set_solution(P.acoustic_state(n, order, c_true))
    truth.set_timing(True)
    t0 = time.perf_counter()
    obs = truth.run(steps)
    wall = time.perf_counter() - t0
    ph = truth.phase_times()
    truth.close()
    ac = make(c0, dt)
    t1 = time.perf_counter()
    J, g = ac.misfit_gradient(steps, obs, checkpoint_every=every)
    grad_wall = time.perf_counter() - t1
    ms_fwd, ms_rec, ms_adj = ac.last_times()
    ac.close()
    # the optimised kernels' split Cauchy-Kowalevski predictor on the same forward run
    sctx = E.Context(3, order, nvar=4, npar=2, pde=E.PDE_ACOUSTIC)
    sctx.set_dialect(E.DIALECT_SPLIT_CK)
    sp = E.Acoustic(sctx, n, 1.0 / n)
    sp.set_material(rho=np.ones(n ** 3), c=c_true)
    sp.set_dt(dt)
    sp.set_source(np.array(P.ACOUSTIC_SOURCE), P.ACOUSTIC_AMP, P.ACOUSTIC_F0, P.ACOUSTIC_T0)
    sp.set_receivers(rec)
    if args.warmup:
        sp.run(args.warmup)
        sp.set_step(0)
        sp.set_solution(P.acoustic_state(n, order, c_true))
    sp.set_timing(True)
    t2 = time.perf_counter()
    obs_split = sp.run(steps)
    split_wall = time.perf_counter() - t2
    sph = sp.phase_times()
    sp.close()
    sctx.close()
    split = {"value": n ** 3 * B ** 3 * 4 * steps / split_wall, "ms_per_step": split_wall * 1e3 / steps,
             "predictor_ms": sph["predictor_ms"] / steps, "corrector_ms": sph["corrector_ms"] / steps,
             "receiver_rel_diff_vs_generic": float(np.linalg.norm(obs_split - obs) / np.linalg.norm(obs))}
    log(f"[bench] acoustic split-CK dialect: {split['value']:.4e} DoF/s, {split['ms_per_step']:.3f} ms/step "
        f"(pred {split['predictor_ms']:.3f} corr {split['corrector_ms']:.3f}); receivers differ from the generic "
        f"dialect by {split['receiver_rel_diff_vs_generic']:.2e}")
    cells = n ** 3
    dofs = cells * B ** 3 * 4 * steps
    pred_b, corr_b = acoustic_bytes_per_cell(B)
    pred_ms, corr_ms = ph["predictor_ms"] / steps, ph["corrector_ms"] / steps
    dom = ("predictor", pred_b, pred_ms) if pred_ms >= corr_ms else ("corrector", corr_b, corr_ms)
    achieved = cells * dom[1] / (dom[2] * 1e-3) / 1e9
    log(f"[bench] acoustic forward: {dofs / wall:.4e} DoF/s, {wall * 1e3 / steps:.3f} ms/step "
        f"(pred {pred_ms:.3f} corr {corr_ms:.3f}); gradient {grad_wall:.2f} s "
        f"(fwd {ms_fwd:.0f} ms, recompute {ms_rec:.0f} ms, adjoint {ms_adj:.0f} ms), J={J:.4e}")
    cpu = None
    if not args.no_cpu:
        O = __import__("oracle")
        threads, visible = cpu_threads()
        cpu = acoustic_cpu_baseline(O, P, order, args.cpu_n, args.cpu_steps, threads)
        log(f"[bench] cpu baseline: {cpu['value']:.4e} DoF/s on {threads} threads")
    out = {
        "metric": "DoF-updates/s 3D linear acoustics order 4 (forward, FP64); misfit-gradient wall time",
        "value": dofs / wall, "unit": "DoF-updates/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D linear acoustics ADER-DG order {order}, {n}^3 periodic cells, c = 1 + 8 seeded "
                               f"Gaussian bumps (+-0.2, width 0.1), rho = 1, Ricker source f0 = 5, 64 receivers, "
                               f"{steps} steps, dt = {dt:.4e}; gradient by the discrete adjoint, checkpoints "
                               f"every {every}", "cells": cells, "order": order},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "phases_ms_per_step": {"predictor": pred_ms, "corrector": corr_ms},
        "predictor_form": ("x-line", "node")[args.acoustic_predictor],
        "gradient_s": grad_wall,
        "gradient_phases_ms": {"forward": ms_fwd, "recompute": ms_rec, "adjoint": ms_adj},
        "gradient_over_forward": grad_wall / wall,
        "misfit": J,
        "split_ck_forward": split,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_batched(args, E, P):
    order, n, steps = 4, 27, args.steps
    B = order + 1
    c3 = E.Context(3, order, 5, 0)
    c2 = E.Context(2, order, 5, 0)
    refined = P.amr_ball_refinement(n, radius=0.25)
    sizes = P.ensemble_sizes(seed=args.seed, count=64)
    m = E.Amr(c3, n, refined)
    level, coords = m.leaves()
    u3 = P.amr_density_wave(order, n, level, coords)
    e = E.Ensemble(c2, sizes)
    u2 = [P.density_wave_2d(int(s), order).ravel() for s in sizes]

    def reset():
        m.set_solution(u3)
        m.init_dt(0.0)
        for i, u in enumerate(u2):
            e.set_solution(i, u)
        e.init_dt(0.0)

    reset()
    t0 = time.perf_counter()
    tune3 = m.autotune()
    tune2 = e.autotune()
    tune_s = time.perf_counter() - t0
    g = E.Group([m], [e])
    if args.warmup:
        g.run(args.warmup, use_graph=True)
    reset()
    t0 = time.perf_counter()
    g.run(steps, use_graph=True)  # synchronous
    wall = time.perf_counter() - t0
    dofs_step = (m.nleaves * B ** 3 + int(np.sum(sizes.astype(np.int64) ** 2)) * B ** 2) * 5
    it3 = executed_iterations(m.picard(), 3, 2 * B, True)
    it2 = executed_iterations(e.picard(), 2, 2 * B, True)
    f3 = float(np.sum(predictor_flops_per_cell(3, B, 5, 5, it3)))
    f2 = float(np.sum(predictor_flops_per_cell(2, B, 5, 5, it2)))
    ach = f3 / (tune3["ms"][0] * 1e-3) / 1e12
    log(f"[bench] batched: {dofs_step * steps / wall:.4e} DoF/s, {wall * 1e3 / steps:.3f} ms/step; "
        f"3D two-level {m.nleaves} leaves (pred {tune3['ms'][0]:.3f} ms, riem {tune3['ms'][1]:.3f}, "
        f"corr {tune3['ms'][2]:.3f}), 2D ensemble {e.ncells()} cells (pred {tune2['ms'][0]:.3f} ms, "
        f"riem {tune2['ms'][1]:.3f}, corr {tune2['ms'][2]:.3f}); tuning {tune_s:.2f} s")
    cpu = None
    if not args.no_cpu:
        O = __import__("oracle")
        threads, visible = cpu_threads()
        o2 = O.Oracle(2, order)
        sample = [int(s) for s in sizes[:32]]
        nst = 4
        t1 = time.perf_counter()
        for s_ in sample:
            luh = np.ascontiguousarray(P.density_wave_2d(s_, order).ravel())
            o2.run_uniform(luh, s_, 1.0 / s_, nst, 0.0, threads=threads)
        o3 = O.AmrOracle(3, order)
        r9 = P.amr_ball_refinement(9, radius=0.25)
        lv, co = O.AmrOracle.amr_leaves(9, r9)
        u = P.amr_density_wave(order, 9, lv, co)
        o3.amr_run(u, 9, r9, nst, threads=threads)
        t = time.perf_counter() - t1
        dofs = nst * (sum(s_ * s_ for s_ in sample) * B ** 2 + len(lv) * B ** 3) * 5
        cpu = dict(value=dofs / t, unit="DoF-updates/s", cores=threads, kind="port",
                   sample=f"oracle on the first 32 2D instances and a 9^3 two-level mesh with the same ball, "
                          f"{nst} steps each ({t:.2f} s wall, {threads} threads)")
        log(f"[bench] cpu baseline: {cpu['value']:.4e} DoF/s on {threads} threads")
    out = {
        "metric": "DoF-updates/s batched uneven work: 3D Euler order-4 two-level mesh + 64 2D Euler instances",
        "value": dofs_step * steps / wall, "unit": "DoF-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D Euler order {order} on 27^3 coarse cells with the ball r=0.25 refined 3x "
                               f"({m.nleaves} leaves, one global dt) + 64 2D Euler order {order} instances of "
                               f"{int(sizes.min())}..{int(sizes.max())} cells per side ({e.ncells()} cells, one dt "
                               f"each); one captured hipGraph per step, autotuned launch geometry",
                   "leaves_3d": int(m.nleaves), "cells_2d": int(e.ncells())},
        "roofline": {"bound": "fp64-valu", "kernel": "3D predictor", "achieved": ach,
                     "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s", "frac": ach / PEAK_TFLOPS["fp64"],
                     "traffic": None},
        "autotune": {"amr": tune3, "ensemble": tune2, "seconds": tune_s},
        "predictor_tflops_2d": f2 / (tune2["ms"][0] * 1e-3) / 1e12,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    g.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("euler", "acoustic", "batched"), default="euler")
    ap.add_argument("--n", type=int, default=None, help="cells per dimension (81 euler, 54 acoustic)")
    ap.add_argument("--order", type=int, default=None, help="order (5 euler, 4 acoustic)")
    ap.add_argument("--picard-fp32", type=int, default=0,
                    help="fixed Picard count of the FP32 form (0: one with the CERK guess and the Picard guard, "
                         "else the largest count the FP64 form executed)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--guess", choices=("trivial", "cerk"), default="cerk",
                    help="initial guess of the FP32 form's Picard loop (with --sweep-variants: of every form)")
    ap.add_argument("--acoustic-predictor", type=int, default=0, choices=(0, 1),
                    help="acoustic forward predictor: 0 x-line kernel, 1 node kernel")
    ap.add_argument("--modes", type=int, default=8)
    ap.add_argument("--slabs", type=int, default=0,
                    help="FP32 one-iteration CERK form: z-plane slabs of the pipelined step (the corrector of a "
                         "slab beside the predictors of the later ones on a second stream); 0: one launch per phase")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-n", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=None)
    ap.add_argument("--sweep-variants", action="store_true",
                    help="only time the predictor kernel variants (node-column, line, line+MFMA, tile+MFMA) and exit")
    ap.add_argument("--variant", type=int, default=None,
                    help="with --sweep-variants: time only this predictor variant (1..4)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (2^27 float64) three times: the reference "
                         "kernel for scaling the FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the engine runs on one GPU (asked for --gpus {args.gpus}); sharding the mesh "
              "over several GPUs is a design, not a runnable mode.", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: launched with WORLD_SIZE > 1; one GPU only", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] copied 3 x 1 GiB")
        return

    import exahype_amd as E
    from exahype_amd import problems as P

    if args.workload == "acoustic":
        args.n = args.n or 54
        args.order = args.order or 4
        args.cpu_n = args.cpu_n or 27
        args.cpu_steps = args.cpu_steps or 20
        if "--steps" not in sys.argv:
            args.steps = 300
        log(f"[bench] acoustic N={args.order} {args.n}^3 cells, steps={args.steps} warmup={args.warmup}")
        bench_acoustic(args, E, P)
        return
    if args.workload == "batched":
        log(f"[bench] batched uneven work, steps={args.steps} warmup={args.warmup}")
        bench_batched(args, E, P)
        return
    args.n = args.n or 81
    args.order = args.order or 5
    args.cpu_n = args.cpu_n or 54   # ~10 s of 16-thread CPU work
    args.cpu_steps = args.cpu_steps or 4

    log(f"[bench] 3D Euler N={args.order} {args.n}^3 cells, steps={args.steps} warmup={args.warmup}")
    if args.sweep_variants:
        k32 = args.picard_fp32 if args.picard_fp32 > 0 else 3
        for prec, name, pic in ((E.PREC_FP64, "fp64", 0), (E.PREC_FP32, "fp32", k32)):
            for variant, vname in ((1, "node-column"), (2, "line"), (3, "line+mfma"), (4, "tile+mfma")):
                if (prec == E.PREC_FP64 and variant in (3, 4)) or \
                        (args.variant is not None and variant != args.variant):
                    continue
                r = run_form(E, P, args.n, args.order, args.steps, args.warmup, prec, pic, args.seed, args.modes, variant,
                             guess=1 if args.guess == "cerk" else 0)
                log(f"[sweep] {name} {vname:12s} predictor {r['predictor_ms']:.2f} ms/step "
                    f"({r['roofline']['achieved']:.2f} TFLOP/s, {r['roofline']['achieved_executed']:.2f} executed), "
                    f"picard {r['picard_hist']}")
        return
    r64 = run_form(E, P, args.n, args.order, args.steps, args.warmup, E.PREC_FP64, 0, args.seed, args.modes)
    log(f"[bench] fp64: {r64['value']:.4e} DoF/s, {r64['ms_per_step']:.2f} ms/step "
        f"(pred {r64['predictor_ms']:.2f} riem {r64['riemann_ms']:.2f} corr {r64['corrector_ms']:.2f}), "
        f"picard {r64['picard_hist']}, {r64['roofline']['achieved']:.2f} TFLOP/s "
        f"({r64['roofline']['achieved_executed']:.2f} executed)")
    cerk = args.guess == "cerk"
    r64c = None
    if cerk:
        # the FP64 form with the FP32 form's initial guess (CERK) and the generic adaptive rule:
        # its largest executed Picard count is the FP32 form's fixed count
        r64c = run_form(E, P, args.n, args.order, args.steps, args.warmup, E.PREC_FP64, 0, args.seed, args.modes,
                        guess=1)
        log(f"[bench] fp64 (CERK guess): {r64c['value']:.4e} DoF/s, {r64c['ms_per_step']:.2f} ms/step "
            f"(pred {r64c['predictor_ms']:.2f} corr {r64c['corrector_ms']:.2f}), picard {r64c['picard_hist']}")
    # FP32 fixed count: one iteration with the CERK guess (cells the Picard guard
    # flags are recomputed with three), else the FP64 form's largest executed count
    k32 = args.picard_fp32 if args.picard_fp32 > 0 else (1 if cerk else r64["picard_max_executed"])
    # slab-pipelined step where it applies (one-iteration CERK form)
    slabs = args.slabs if (cerk and k32 == 1 and args.slabs >= 3) else 0
    r32 = run_form(E, P, args.n, args.order, args.steps, args.warmup, E.PREC_FP32, k32, args.seed, args.modes,
                   guess=1 if cerk else 0, slabs=slabs)
    if r32["one_launch_step"]:
        log(f"[bench] fp32 one launch per phase: {r32['one_launch_step']['ms_per_step']:.2f} ms/step, "
            f"bitwise equal to the {r32['slabs']}-slab step: {r32['one_launch_step']['bitwise_equal']}")
    log(f"[bench] fp32: {r32['value']:.4e} DoF/s, {r32['ms_per_step']:.2f} ms/step "
        f"(pred {r32['predictor_ms']:.2f} riem {r32['riemann_ms']:.2f} corr {r32['corrector_ms']:.2f}), "
        f"guess {args.guess}, picard {k32}, guard recomputed {r32['guard_recomputed_cells']} cells, "
        f"{r32['roofline']['achieved']:.2f} TFLOP/s")

    def rel_err(x, y):
        a = x.reshape(-1, 5)
        b = y.reshape(-1, 5)
        return [float(np.linalg.norm(a[:, v] - b[:, v]) / np.linalg.norm(b[:, v])) for v in range(5)]

    # errors against the FP64 form of the reference's generic kernels (trivial guess, adaptive Picard)
    err = rel_err(r32.pop("solution"), r64["solution"])
    log(f"[bench] rel L2 error fp32 vs fp64 per variable: {['%.2e' % e for e in err]}")
    r32t = None
    if cerk:
        k32t = r64["picard_max_executed"]
        r32t = run_form(E, P, args.n, args.order, args.steps, args.warmup, E.PREC_FP32, k32t, args.seed, args.modes)
        r32t["rel_l2_err_vs_fp64"] = rel_err(r32t.pop("solution"), r64["solution"])
        r32t["picard_fixed"] = k32t
        log(f"[bench] fp32 (trivial guess): {r32t['value']:.4e} DoF/s, {r32t['ms_per_step']:.2f} ms/step "
            f"(pred {r32t['predictor_ms']:.2f}), picard {k32t}, max err {max(r32t['rel_l2_err_vs_fp64']):.2e}")
        r64c["rel_l2_err_vs_fp64"] = rel_err(r64c.pop("solution"), r64["solution"])
    # BF16 experiment: the tile kernel with BF16 matrix-core operands, beside the same kernel in FP32,
    # both with the trivial guess and the FP64 form's executed Picard count
    k16 = r64["picard_max_executed"]
    r16 = run_form(E, P, args.n, args.order, args.steps, args.warmup, E.PREC_BF16, k16, args.seed, args.modes,
                   variant=4)
    r16["rel_l2_err_vs_fp64"] = rel_err(r16.pop("solution"), r64["solution"])
    r32tile = run_form(E, P, args.n, args.order, args.steps, args.warmup, E.PREC_FP32, k16, args.seed, args.modes,
                       variant=4)
    r32tile["rel_l2_err_vs_fp64"] = rel_err(r32tile.pop("solution"), r64["solution"])
    log(f"[bench] bf16 tile: {r16['value']:.4e} DoF/s, pred {r16['predictor_ms']:.2f} ms, picard {k16}, "
        f"err {['%.2e' % e for e in r16['rel_l2_err_vs_fp64']]}; fp32 tile: pred {r32tile['predictor_ms']:.2f} ms, "
        f"max err {max(r32tile['rel_l2_err_vs_fp64']):.2e}")
    r64.pop("solution")

    cpu = None
    if not args.no_cpu:
        O = __import__("oracle")
        threads, visible = cpu_threads()
        cpu = cpu_baseline(O, P, args.order, args.cpu_n, args.cpu_steps, threads, visible)
        log(f"[bench] cpu baseline: {cpu['value']:.4e} DoF/s on {threads} threads")

    out = {
        "metric": "DoF-updates/s 3D Euler order 5: FP32-MFMA vs FP64 at <=1e-6 rel L2 error",
        "value": r32["value"],
        "unit": "DoF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"3D compressible Euler ADER-DG order {args.order} ({args.order + 1}^3 nodes), 5 vars, "
                        f"{args.n}^3 periodic cells, density wave + {args.modes} seeded Fourier modes "
                        f"(amp 1e-2, |k|<=4) in rho and u, CFL 0.9, {args.steps} global steps; "
                        + (f"FP32 predictor with the CERK initial guess of the reference's optimised kernels and "
                           f"{k32} fixed Picard iteration(s); the Picard guard recomputes a cell with 3 iterations "
                           f"when the first iteration changes its time-averaged predictor by >= 1e-6 relative"
                           if cerk else
                           f"FP32 predictor (fixed {k32} Picard iterations)")
                        + ", FP64 Riemann/corrector/update/CFL; errors vs the FP64 generic kernels (trivial guess, "
                          "adaptive Picard)",
            "cells": args.n ** 3, "order": args.order, "picard_fp32": k32, "guess_fp32": args.guess,
            "seed": args.seed, "slabs": r32["slabs"],
        },
        "roofline": r32["roofline"],
        "cpu_baseline": cpu,
        "fp64": {k: r64[k] for k in ("value", "ms_per_step", "predictor_ms", "riemann_ms", "corrector_ms",
                                     "picard_hist", "roofline")},
        "fp32_phases_ms": {k: r32[k] for k in ("predictor_ms", "riemann_ms", "corrector_ms")},
        "step_form": "two kernels per step: predictor (+ volume part of the update), then one corrector "
                     "kernel (Riemann problems of every cell's faces + face integrals + update + CFL); "
                     "riemann_ms is the boundary-condition slot between them (periodic: empty)"
                     + (f"; timed in {r32['slabs']} slabs of z planes, the corrector of a slab on a second stream "
                        f"beside the predictors of the later slabs (fp32_phases_ms and the roofline: the same steps "
                        f"re-run with one launch per phase and HIP events, fp32_one_launch_step)"
                        if r32["slabs"] else ""),
        "fp32_one_launch_step": r32["one_launch_step"],
        "fp32_guard_recomputed_cells": r32["guard_recomputed_cells"],
        "speedup_fp32_vs_fp64": r32["value"] / r64["value"],
        "rel_l2_err_fp32_vs_fp64": err,
        "max_rel_l2_err": max(err),
    }
    if cerk:
        keys = ("value", "ms_per_step", "predictor_ms", "corrector_ms", "picard_hist", "rel_l2_err_vs_fp64")
        out["fp64_cerk_guess"] = {k: r64c[k] for k in keys}
        out["fp32_trivial_guess"] = {k: r32t[k] for k in keys if k != "picard_hist"}
        out["fp32_trivial_guess"]["picard_fixed"] = r32t["picard_fixed"]
        out["speedup_fp32_vs_fp64_same_guess"] = r32["value"] / r64c["value"]
    keys = ("value", "ms_per_step", "predictor_ms", "corrector_ms", "rel_l2_err_vs_fp64")
    out["bf16_experiment"] = dict({k: r16[k] for k in keys}, picard_fixed=k16, guess="trivial",
                                  kernel="predictor3d_tile_kernel<6, BF16 operands>",
                                  predictor_tflops=r16["roofline"]["achieved"],
                                  fp32_same_kernel={k: r32tile[k] for k in keys})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
