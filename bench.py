#!/usr/bin/env python3
"""Benchmark of the hesthaven DGTD operator + fused LSERK4 stage on one MI355X.

Workload (a "step" is one LSERK4 time step = 5 fused stage launches over the
whole mesh): PEC cavity [0,1]^3, 32^3 hexahedra x 6 = 196,608 tetrahedra,
order 4 (Np = 35, 41,287,680 DOF), initial fields a seeded Gaussian pulse
(centre uniform in [0.3,0.7]^3, width 0.05, amplitude 1 on each of the six
components), dt from the Hesthaven rule dt = 1/(max fscale * p^2).

The headline value is FP32 DOF*RK-stages/s; the FP64 path runs the same steps
beside it and the relative L2 difference FP32 vs FP64 is reported.  roofline
takes the algorithmic bytes per element per stage (24*Np*s + 27*s + 32) and
flops (36 Np^2 + 48 Np Nfp + 240 Nfp + 96 Np) times K over the average
stage-kernel duration from HIP events on the operator's stream, and reports
the resource whose ideal time is the longer (at order 4 in FP32 the matrix
cores: 95 us against 86 us of HBM traffic) with the HBM figures beside it.  cpu_baseline times the CPU FP64 oracle (a restatement of the
reference algorithm) with OpenMP on the host cores for one LSERK4 step of the
same mesh.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def gaussian_all_components(pos, seed, spread=0.05):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.3, 0.7, 3)
    r2 = (pos[0] - c[0]) ** 2 + (pos[1] - c[1]) ** 2 + (pos[2] - c[2]) ** 2
    g = np.exp(-r2 / (2.0 * spread ** 2))
    return np.tile(g.ravel(), 6), c


def time_path(D, vx, E, fbc, order, precision, q0, dt, steps, warmup, kernel="auto"):
    op = D.HesthavenEvolution(vx, E, fbc, order, precision=precision, integrator="lserk4",
                              kernel=kernel)
    variant = op.kernel_variant()
    kname, tuning = stage_kernel(op, order, precision)
    op.set_state(q0, 0.0)
    if warmup:
        op.advance(dt, warmup)
    op.set_timing(True)
    t0 = time.perf_counter()
    op.advance(dt, steps)  # returns after the operator's stream has drained
    wall = time.perf_counter() - t0
    kern_us, launches = op.last_kernel_us()
    state = op.get_state()
    nm = op.norms()
    bytes_el, flops_el = op.work_per_element()
    op.close()
    return dict(wall=wall, kern_us=kern_us, launches=launches, state=state, norms=nm,
                bytes_el=bytes_el, flops_el=flops_el, variant=variant, kernel=kname, tuning=tuning)


def cpu_baseline(vx, E, fbc, order, q0, dt, budget_s=10.0, threads=None, eps=None):
    """LSERK4 steps (5 operator applications each) of the oracle on the host,
    as many as fit in about budget_s seconds (at least one), on `threads`
    OpenMP threads (default: all)."""
    from oracle import oracle as O

    ref = O.Operator(vx, E.reshape(-1, 4), fbc.reshape(-1, 4).astype(np.int32), order, eps=eps)
    all_threads = O.threads()
    if threads:
        O.set_threads(threads)
    try:
        q = q0.copy()
        res = np.zeros_like(q)
        steps = 0
        t0 = time.perf_counter()
        while True:
            ref.lserk4_step(q, res, steps * dt, dt)
            steps += 1
            wall = time.perf_counter() - t0
            if wall >= budget_s or wall / steps * (steps + 1) > 2.0 * budget_s:
                break
        used = O.threads()
    finally:
        O.set_threads(all_threads)
    return wall, steps, used


def profiled_traffic(kernel, workload):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (calibrated against a plain copy) for this kernel and workload, from the
    counter summaries under profiles/; None when no summary matches."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None, None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith("_hbm_counters.md"):
            continue
        # records of "kernel = ...", "workload = ...", "traffic_bytes_per_launch = ..." lines
        kv = {}
        for ln in open(os.path.join(pdir, name)):
            if " = " not in ln:
                continue
            k, v = ln.strip().split(" = ", 1)
            if k == "kernel":
                kv = {}
            kv[k] = v
            if kv.get("kernel") == kernel and kv.get("workload") == workload and "traffic_bytes_per_launch" in kv:
                return float(kv["traffic_bytes_per_launch"]), "profiles/" + name
    return None, None


def mass_norm(D, vx, E, fbc, order, v):
    """sqrt(sum_e J_e v_e^T M v_e) over all six components, from the library's
    FP64 energy reduction (vacuum: energy = norm^2 / 2)."""
    op = D.HesthavenEvolution(vx, E, fbc, order, precision=64, kernel="valu")
    try:
        op.set_state(v, 0.0)
        return math.sqrt(2.0 * op.norms()["energy"])
    finally:
        op.close()


def work_per_element(order, precision):
    """Algorithmic bytes and flops per tetrahedron per stage (the library's
    dgtd_op_work_per_element formula): 24 Np s + 27 s + 32 bytes and
    36 Np^2 + 48 Np Nfp + 240 Nfp + 96 Np flops."""
    s = 4 if precision == 32 else 8
    Np = (order + 1) * (order + 2) * (order + 3) // 6
    Nfp = (order + 1) * (order + 2) // 2
    return 24 * Np * s + 27 * s + 32, 36 * Np * Np + 48 * Np * Nfp + 240 * Nfp + 96 * Np


def stage_kernel(op, order, precision, dim=3):
    """The stage kernel an operator runs, named as in the rocprof summaries
    (matrix-core kernels with their tuning configuration), and the
    autotuner's record for it."""
    t = op.tuning()
    ty = "float" if precision == 32 else "double"
    if t["variant"] == "mfma":
        name = f"dg_stage_mfma<{order},{ty},LSERK,{dim}D,cfg{t['config']}>"
    elif dim == 2:
        name = f"dg_tri_stage<{order},{ty},LSERK>"
    else:
        name = f"dg_stage_kernel<{order},{ty},LSERK>"
    return name, t


def roofline_fracs(bytes_, flops, seconds, precision):
    """HBM and matrix-core fractions of `bytes_` and `flops` done in `seconds`,
    the attainable (max of the two ideal times) fraction, and the bound."""
    peak_tf = FP32_PEAK_TFLOPS if precision == 32 else FP64_PEAK_TFLOPS
    gbs = bytes_ / seconds / 1e9
    tf = flops / seconds / 1e12
    t_hbm = bytes_ / (HBM_PEAK_GBS * 1e9)
    t_mfma = flops / (peak_tf * 1e12)
    return {"bound": "mfma" if t_mfma > t_hbm else "hbm",
            "hbm_achieved_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
            "achieved_tflops": tf, "mfma_frac": tf / peak_tf, "peak_tflops": peak_tf,
            "ideal_us": {"hbm": t_hbm * 1e6, "mfma": t_mfma * 1e6},
            "attainable_frac": max(t_hbm, t_mfma) / seconds}


def te110_exact(pos, t):
    """Analytic TE110 mode of the PEC unit cube (eps = mu = 1) at the nodes."""
    x, y = pos[0], pos[1]
    w = np.pi * np.sqrt(2.0)
    q = np.zeros((6,) + x.shape)
    q[2] = np.sin(np.pi * x) * np.sin(np.pi * y) * np.cos(w * t)
    q[3] = -(np.pi / w) * np.sin(np.pi * x) * np.cos(np.pi * y) * np.sin(w * t)
    q[4] = (np.pi / w) * np.cos(np.pi * x) * np.sin(np.pi * y) * np.sin(w * t)
    return q.ravel()


def bench_c1(D, args, cpu=True):
    """8^3 hexes x 6 = 3,072 tets, p=3, PEC unit cube, TE110 initial mode
    (Ez = sin(pi x) sin(pi y)), upwind, LSERK4 dt = 2e-3 to T = 1: FP64 on
    the GPU end to end against the analytic mode, and the same run of the
    CPU FP64 oracle end to end (all host threads)."""
    n, order, dt, T = 8, 3, 2e-3, 1.0
    vx, E, fbc = D.cube_mesh(n, 1.0, "pec")
    K = E.shape[0]
    x0 = D.project_initial_fields(vx, E, order, [D.resonant((1, 1), "E", (0, 0, 1))])
    dof = 6 * D.num_nodes(order) * K
    op = D.HesthavenEvolution(vx, E, fbc, order, precision=64, integrator="lserk4")
    pos = op.node_positions()
    op.set_state(x0)
    op.advance(dt, 2)
    op.set_state(x0)
    t0 = time.perf_counter()
    steps = op.run(dt, T)
    wall = time.perf_counter() - t0
    xg = op.get_state()
    variant = op.kernel_variant()
    op.close()
    ex = te110_exact(pos, T)
    out = {"workload": f"PEC cavity [0,1]^3, {n}^3 hexes x 6 = {K} tets, p={order}, TE110, LSERK4 dt={dt}, T={T}",
           "dof": dof, "steps": steps,
           "fp64": {"value": dof * 5 * steps / wall, "unit": "DOF*RK-stages/s", "wall_s": wall,
                    "kernel_variant": variant, "us_per_step": 1e6 * wall / steps},
           "rel_l2_vs_analytic": float(np.linalg.norm(xg - ex) / np.linalg.norm(ex))}
    if cpu:
        from oracle import oracle as O

        ref = O.Operator(vx, E.reshape(-1, 4), fbc.reshape(-1, 4).astype(np.int32), order)
        x = x0.copy()
        t0 = time.perf_counter()
        csteps, _ = ref.run(x, dt, T, "lserk4")
        cwall = time.perf_counter() - t0
        out["rel_l2_gpu_fp64_vs_oracle"] = float(np.linalg.norm(xg - x) / np.linalg.norm(x))
        out["cpu_baseline"] = {"value": dof * 5 * csteps / cwall, "unit": "DOF*RK-stages/s",
                               "cores": O.threads(), "kind": "port", "wall_s": cwall,
                               "sample": f"the whole run ({csteps} LSERK4 steps) of the FP64 CPU oracle, "
                                         f"OpenMP {O.threads()} threads"}
    return out


def bench_c3(D, args):
    """48^3 hexes x 6 = 663,552 tets on [0,2]^3, p=3, eps_r ~ U[1,4] per element
    (seeded), mu_r = 1, SMA on all sides, Ez Gaussian pulse, dt = 5e-4,
    FP32 state with FP64 energy reductions, 2000 LSERK4 steps (FP64 beside it)."""
    n, L, order, dt = 48, 2.0, 3, 5e-4
    steps = args.steps if args.workload == "c3" and args.steps != 100 else 2000
    vx, E, fbc = D.cube_mesh(n, L, "sma")
    K = E.shape[0]
    rng = np.random.default_rng(args.seed)
    eps = rng.uniform(1.0, 4.0, K)
    c = rng.uniform(0.3, 0.7, 3) * L
    x0 = D.project_initial_fields(vx, E, order, [D.gaussian("E", c, 0.05 * L, (0, 0, 1))])
    dof = 6 * D.num_nodes(order) * K
    out = {"workload": f"SMA box [0,{L}]^3, {n}^3 hexes x 6 = {K} tets, p={order}, eps_r~U[1,4], "
                       f"Ez Gaussian at ({c[0]:.3f},{c[1]:.3f},{c[2]:.3f}), dt={dt}, {steps} LSERK4 steps",
           "dof": dof}
    states = {}
    for prec in (32, 64):
        op = D.HesthavenEvolution(vx, E, fbc, order, eps=eps, precision=prec, integrator="lserk4",
                                  kernel=args.kernel)
        op.set_state(x0)
        e0 = op.norms()["energy"]
        op.advance(dt, args.warmup)
        op.set_timing(True)
        t0 = time.perf_counter()
        op.advance(dt, steps)
        wall = time.perf_counter() - t0
        us, launches = op.last_kernel_us()
        nm = op.norms()
        b, f = op.work_per_element()
        states[prec] = op.get_state()
        var = op.kernel_variant()
        kname, tuning = stage_kernel(op, order, prec)
        out[f"fp{prec}"] = {"value": dof * 5 * steps / wall, "unit": "DOF*RK-stages/s", "wall_s": wall,
                            "kernel_avg_us": us, "kernel_variant": var, "kernel": kname, "tuning": tuning,
                            "algorithmic_bytes_per_launch": b * K, "algorithmic_flops_per_launch": f * K,
                            "roofline": roofline_fracs(b * K, f * K, us * 1e-6, prec),
                            "hbm_frac": b * K / (us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                            "energy_initial": e0, "energy_final": nm["energy"], "nonfinite": nm["nonfinite"]}
        if prec == 32:
            traffic, src = profiled_traffic(out["fp32"]["kernel"], f"smabox{n}_p{order}_fp32_lserk4")
            out["fp32"]["roofline"]["traffic"] = traffic
            out["fp32"]["roofline"]["traffic_source"] = src
        op.close()
    out["rel_l2_fp32_vs_fp64"] = float(np.linalg.norm(states[32] - states[64]) / np.linalg.norm(states[64]))
    out["fp32_over_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    if not args.no_cpu:
        wall, csteps, threads = cpu_baseline(vx, E, fbc, order, x0, dt, budget_s=8.0, eps=eps)
        out["cpu_baseline"] = {"value": dof * 5 * csteps / wall, "unit": "DOF*RK-stages/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{csteps} LSERK4 step(s) of the same mesh (SMA, random eps), FP64 CPU "
                                         f"oracle, OpenMP {threads} threads, {wall:.2f} s"}
    return out


def bench_global(D, args, shape="tet"):
    """The assembled (Global) operator with conductivity: PEC cavity [0,1]^3,
    12^3 hexes x 6 = 10,368 tets, p=3, eps_r ~ U[1,4], a conducting slab
    (sigma = 2 for x < 1/2), Ez Gaussian pulse, 100 LSERK4 steps in FP32 and
    FP64 through the fused CSR SpMV stage kernel.  shape "quad": the same on
    a PEC/PMC square [0,1]^2 of 256^2 quadrilaterals; "hex": on a PEC cavity
    [0,1]^3 of 32^3 hexahedra (the reference runs its quadrilateral and
    hexahedral meshes on this operator), p=3; --size overrides the cells per
    edge.  CPU leg: the same FP64 CSR product on one host core (scipy), the
    Global operator's SparseMatrix::Mult."""
    order = 3
    steps = args.steps
    rng = np.random.default_rng(args.seed)
    quads = shape == "quad"
    if quads:
        n = args.size or 256
        vx, E, fbc = D.cartesian_2d(n, n, "quad", bc=("pmc", "pec", "pmc", "pec"))
        np_ = (order + 1) ** 2
        src = [D.gaussian("E", (0.6, 0.45, 0.0), 0.1, (0, 0, 1), dimension=2)]
        mesh = f"PEC/PMC square [0,1]^2, {n}^2 quadrilaterals"
    elif shape == "hex":
        n = args.size or 32
        vx, E, fbc = D.cartesian_3d(n, n, n, "hex", bc="pec")
        np_ = (order + 1) ** 3
        src = [D.gaussian("E", (0.6, 0.45, 0.55), 0.1, (0, 0, 1))]
        mesh = f"PEC cavity [0,1]^3, {n}^3 hexahedra"
    else:
        n = args.size or 12
        vx, E, fbc = D.cube_mesh(n, 1.0, "pec")
        np_ = D.num_nodes(order)
        src = [D.gaussian("E", (0.6, 0.45, 0.55), 0.1, (0, 0, 1))]
        mesh = f"PEC cavity [0,1]^3, {n}^3 hexes x 6 = {E.shape[0]} tets"
    K = E.shape[0]
    eps = rng.uniform(1.0, 4.0, K)
    cx = vx.reshape(-1, 3)[E].mean(axis=1)[:, 0]
    sigma = np.where(cx < 0.5, 2.0, 0.0)
    x0 = D.project_initial_fields(vx, E, order, src)
    dof = 6 * np_ * K
    out = {"workload": f"{mesh}, p={order}, eps_r~U[1,4], "
                       f"sigma=2 for x<1/2, Ez Gaussian, {steps} LSERK4 steps, assembled CSR operator", "dof": dof}
    states = {}

    def timed(op, prec, name):
        dt = 0.5 / (op.max_fscale() * order ** 2)
        op.set_state(x0)
        e0 = op.norms()["energy"]
        op.advance(dt, args.warmup)
        op.set_timing(True)
        t0 = time.perf_counter()
        op.advance(dt, steps)
        wall = time.perf_counter() - t0
        us, launches = op.last_kernel_us()
        nm = op.norms()
        b, f = op.work_per_element()
        state = op.get_state()
        op.close()
        return state, {"value": dof * 5 * steps / wall, "unit": "DOF*RK-stages/s", "wall_s": wall,
                       "kernel_avg_us": us, "kernel": f"{name}<{'float' if prec == 32 else 'double'},LSERK>",
                       "algorithmic_bytes_per_launch": b * K,
                       "hbm_frac": b * K / (us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                       "energy_initial": e0, "energy_final": nm["energy"], "nonfinite": nm["nonfinite"]}

    for prec in (32, 64):
        op = D.GlobalEvolution(vx, E, fbc, order, eps=eps, sigma=sigma, precision=prec, integrator="lserk4")
        states[prec], out[f"fp{prec}"] = timed(op, prec, "csr_stage")
    if shape == "hex":  # the matrix-free hexahedron stage on the same operator
        for prec in (32, 64):
            op = D.HesthavenEvolution(vx, E, fbc, order, eps=eps, sigma=sigma, precision=prec,
                                      integrator="lserk4", kernel="valu")
            st, out[f"matrix_free_fp{prec}"] = timed(op, prec, "hex_stage")
            out[f"matrix_free_fp{prec}"]["rel_l2_vs_csr"] = float(np.linalg.norm(st - states[prec])
                                                                   / np.linalg.norm(states[prec]))
    out["rel_l2_fp32_vs_fp64"] = float(np.linalg.norm(states[32] - states[64]) / np.linalg.norm(states[64]))
    if not args.no_cpu:
        import scipy.sparse as sp

        rowptr, col, val = D.global_assemble(vx, E, fbc, order, eps=eps, sigma=sigma, quads=quads or None)
        A = sp.csr_matrix((val, col, rowptr), shape=(dof, dof))
        out["nnz"] = int(val.size)
        x = x0.copy()
        A @ x
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 10.0:
            x = A @ x
            reps += 1
        wall = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": dof * reps / wall, "unit": "DOF*RK-stages/s", "cores": 1, "kind": "port",
                               "sample": f"{reps} FP64 CSR products (scipy) of the same matrix, the stage's SpMV only"}
    print(json.dumps(out))


def bench_tri(D, args):
    """The 2D operator: PEC square [0,1]^2, 512^2 squares x 2 = 524,288
    triangles, p=4 (Np=15), Ez Gaussian pulse, dt from the Hesthaven rule
    (cfl 0.5), 200 LSERK4 steps in FP32 and FP64."""
    n, order = 512, 4
    steps = args.steps if args.steps != 100 else 200
    vx, E, fbc = D.square_mesh(n, 1.0, "pec")
    K = E.shape[0]
    Np = (order + 1) * (order + 2) // 2
    c = (0.43, 0.57, 0.0)
    x0 = D.project_initial_fields(vx, E, order, [D.gaussian("E", c, 0.05, (0, 0, 1), dimension=2)])
    dof = 6 * Np * K
    out = {"workload": f"PEC square [0,1]^2, {n}^2 squares x 2 = {K} triangles, p={order} (Np={Np}), "
                       f"{dof} DOF, Ez Gaussian at ({c[0]},{c[1]}), {steps} LSERK4 steps", "dof": dof}
    states = {}
    for prec in (32, 64):
        op = D.HesthavenEvolution(vx, E, fbc, order, precision=prec, integrator="lserk4", kernel=args.kernel)
        dt = 0.5 / (op.max_fscale() * order ** 2)
        op.set_state(x0)
        e0 = op.norms()["energy"]
        op.advance(dt, args.warmup)
        op.set_timing(True)
        t0 = time.perf_counter()
        op.advance(dt, steps)
        wall = time.perf_counter() - t0
        us, launches = op.last_kernel_us()
        nm = op.norms()
        b, f = op.work_per_element()
        states[prec] = op.get_state()
        kname, tuning = stage_kernel(op, order, prec, dim=2)
        out[f"fp{prec}"] = {"value": dof * 5 * steps / wall, "unit": "DOF*RK-stages/s", "wall_s": wall,
                            "kernel_avg_us": us, "kernel_variant": op.kernel_variant(),
                            "kernel": kname, "tuning": tuning,
                            "algorithmic_bytes_per_launch": b * K,
                            "hbm_frac": b * K / (us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                            "achieved_tflops": f * K / (us * 1e-6) / 1e12,
                            "energy_initial": e0, "energy_final": nm["energy"], "nonfinite": nm["nonfinite"]}
        op.close()
    out["rel_l2_fp32_vs_fp64"] = float(np.linalg.norm(states[32] - states[64]) / np.linalg.norm(states[64]))
    out["fp32_over_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    print(json.dumps(out))


def bench_c4(D, args):
    """64 independent PEC cavities, edge n_i log-uniform over {2..24} hexes,
    order uniform over {1..5}, seeded Gaussian pulses, FP32, one persistent
    launch per stage, hipGraph-replayed steps; all run to final time 0.1."""
    from dgtd_amd import workloads as W

    T = 0.1
    insts = W.cavity_batch(args.seed, 64)
    res = {}
    for prec in (32, 64):
        b = D.Batch(precision=prec, max_instances=64)
        ids = [b.add(vx, E, fbc, order, x0=x0, final_time=T) for (n, order, vx, E, fbc, x0) in insts]
        b.poll()
        st = b.stats()
        t0 = time.perf_counter()
        launched = b.run(poll_every=16)
        wall = time.perf_counter() - t0
        work = 0
        bytes_, flops = 0.0, 0.0
        states = []
        for (n, order, vx, E, fbc, x0), i in zip(insts, ids):
            info = b.info(i)
            K = E.shape[0]
            stages = 5 * info["steps"]
            work += 6 * D.num_nodes(order) * K * stages
            be, fe = work_per_element(order, prec)
            bytes_ += be * K * stages
            flops += fe * K * stages
            states.append(b.get_state(i))
        res[prec] = dict(value=work / wall, wall_s=wall, graph_steps=launched, items=st["items"],
                         grid=st["grid"], states=states, roofline=roofline_fracs(bytes_, flops, wall, prec),
                         bytes=bytes_, flops=flops)
        b.close()
    errs = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(res[32]["states"], res[64]["states"])]
    out = {"workload": "64 PEC cavities, n~logU{2..24}, p~U{1..5}, Gaussian pulses, T=0.1, LSERK4",
           "instances": [(n, o, int(E.shape[0])) for (n, o, vx, E, fbc, x0) in insts],
           "total_dof": int(sum(6 * D.num_nodes(o) * E.shape[0] for (n, o, vx, E, fbc, x0) in insts))}
    for prec in (32, 64):
        r = res[prec]
        out[f"fp{prec}"] = {"value": r["value"], "unit": "DOF*RK-stages/s", "wall_s": r["wall_s"],
                            "graph_steps": r["graph_steps"], "work_items": r["items"], "grid": r["grid"],
                            "algorithmic_bytes": r["bytes"], "algorithmic_flops": r["flops"],
                            "roofline": r["roofline"]}
    out["max_rel_l2_fp32_vs_fp64"] = max(errs)
    out["fp32_over_fp64"] = res[32]["value"] / res[64]["value"]
    return out


def bench_c5(D, args):
    """Poisson arrivals of cavity / SMA-box solves, K log-uniform 1k..400k tets,
    p in {2,3,4}, final time U[0.5,2], eps_r ~ U[1,4] per element; FP32."""
    from dgtd_amd import serve as S

    reqs = S.poisson_stream(args.rate, args.duration, seed=args.seed)
    srv = S.Server(precision=32, hbm_budget=int(64e9), max_instances=512, poll_every=4, workers=8)
    recs = srv.serve(reqs, max_wall=args.duration * 3 + 60)
    srv.close()
    summ = S.summarize(recs, srv.wall)
    done = [r for r in recs.values() if r.finished > 0]
    busy = sum(r.dof * 5 * r.steps for r in done)
    summ.update({"workload": f"Poisson {args.rate}/s for {args.duration}s, K~logU[1k,400k], p in {{2,3,4}}, "
                             f"T~U[0.5,2], eps_r~U[1,4], FP32",
                 "offered_dof_stages": float(sum(r.dof * 5 * math.ceil(r.final_time * 2 * math.sqrt(2) * r.n
                                                                         * r.order ** 2) for r in reqs))})
    if done:
        summ["sustainable_rate_estimate_per_s"] = len(done) / srv.wall
    print(json.dumps(summ))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=32, help="hexahedra per cube edge")
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison path")
    ap.add_argument("--kernel", default="auto", choices=["auto", "valu", "mfma", "csr"])
    ap.add_argument("--headline-only", action="store_true",
                    help="c2: skip the c1, c3 and c4 sections of the line")
    ap.add_argument("--workload", default="c2",
                    choices=["c1", "c2", "c3", "c4", "c5", "tri", "global", "quad", "hex"],
                    help="c2 (default, the headline): 32^3 cavity p=4; c3: 48^3 SMA box with random eps, "
                         "p=3, 2000 steps; c4: 64 uneven cavities batched; c5: Poisson serving stream; "
                         "tri: the 2D operator on 524,288 triangles, p=4; global: the assembled CSR operator "
                         "on 10,368 tets with conductivity; quad: the same on 256^2 quadrilaterals; hex: on "
                         "32^3 hexahedra")
    ap.add_argument("--size", type=int, default=0, help="global/quad/hex: cells per edge (0 = the workload's own)")
    ap.add_argument("--rate", type=float, default=40.0, help="c5: arrivals per second")
    ap.add_argument("--duration", type=float, default=60.0, help="c5: seconds of arrivals")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (HBM-counter calibration)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch

        src = torch.ones(args.copy_calib, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes": args.copy_calib}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not an implementation)"}))
        sys.exit(2)

    import dgtd_amd as D

    if args.workload == "c1":
        return print(json.dumps(bench_c1(D, args, cpu=not args.no_cpu)))
    if args.workload == "c3":
        return print(json.dumps(bench_c3(D, args)))
    if args.workload == "c4":
        return print(json.dumps(bench_c4(D, args)))
    if args.workload == "c5":
        return bench_c5(D, args)
    if args.workload == "tri":
        return bench_tri(D, args)
    if args.workload == "global":
        return bench_global(D, args)
    if args.workload in ("quad", "hex"):
        return bench_global(D, args, shape=args.workload)
    n, order = args.n, args.order
    vx, E, fbc = D.cube_mesh(n, 1.0, "pec")
    K = E.shape[0]
    Np = D.num_nodes(order)
    dof = 6 * Np * K
    probe = D.HesthavenEvolution(vx, E, fbc, order, precision=64, kernel="valu")  # positions and dt only: no tuning
    pos = probe.node_positions()
    dt = D.estimate_time_step(probe)
    probe.close()
    q0, centre = gaussian_all_components(pos, args.seed)

    r32 = time_path(D, vx, E, fbc, order, 32, q0, dt, args.steps, args.warmup, args.kernel)
    stages = 5 * args.steps
    value = dof * stages / r32["wall"]
    kern_s = r32["kern_us"] * 1e-6
    achieved_gbs = r32["bytes_el"] * K / kern_s / 1e9
    achieved_tf = r32["flops_el"] * K / kern_s / 1e12
    t_hbm = r32["bytes_el"] * K / (HBM_PEAK_GBS * 1e9)
    t_mfma = r32["flops_el"] * K / (FP32_PEAK_TFLOPS * 1e12)
    t_star = max(t_hbm, t_mfma)
    # the bound is the resource whose ideal time is the longer one
    mfma_bound = t_mfma > t_hbm
    kname = r32["kernel"]
    traffic, traffic_src = profiled_traffic(kname, f"cavity{n}_p{order}_fp32_lserk4")
    out = {
        "metric": "DOF*RK-stages/s (FP32 fused LSERK4 stage, hesthaven DGTD operator)",
        "value": value,
        "unit": "DOF*RK-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r32["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian pulse, structured PEC cavity)",
        "config": {
            "workload": f"PEC cavity [0,1]^3, {n}^3 hexes x 6 = {K} tets, order {order} (Np={Np}), "
                        f"{dof} DOF, LSERK4 dt={dt:.6e}, Gaussian width 0.05 centre "
                        f"({centre[0]:.4f},{centre[1]:.4f},{centre[2]:.4f}) on all 6 components",
            "K": K, "order": order, "dof": dof, "dt": dt, "integrator": "lserk4",
            "stages_per_step": 5,
        },
        "roofline": {
            "bound": "mfma" if mfma_bound else "hbm",
            "achieved": achieved_tf if mfma_bound else achieved_gbs,
            "peak": FP32_PEAK_TFLOPS if mfma_bound else HBM_PEAK_GBS,
            "unit": "TFLOP/s" if mfma_bound else "GB/s",
            "frac": achieved_tf / FP32_PEAK_TFLOPS if mfma_bound else achieved_gbs / HBM_PEAK_GBS,
            "hbm_achieved_gbs": achieved_gbs,
            "hbm_frac": achieved_gbs / HBM_PEAK_GBS,
            "ideal_us": {"hbm": t_hbm * 1e6, "mfma": t_mfma * 1e6},
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": kname,
            "kernel_avg_us": r32["kern_us"],
            "launches": r32["launches"],
            "algorithmic_bytes_per_launch": r32["bytes_el"] * K,
            "algorithmic_flops_per_launch": r32["flops_el"] * K,
            "achieved_tflops": achieved_tf,
            "fp32_peak_tflops": FP32_PEAK_TFLOPS,
            "attainable_frac": t_star / kern_s,
            "tuning": r32["tuning"],
        },
    }
    if not args.no_fp64:
        r64 = time_path(D, vx, E, fbc, order, 64, q0, dt, args.steps, args.warmup, args.kernel)
        v64 = dof * stages / r64["wall"]
        k64 = r64["kern_us"] * 1e-6
        out["fp64"] = {
            "value": v64,
            "ms_per_step": 1e3 * r64["wall"] / args.steps,
            "kernel_avg_us": r64["kern_us"],
            "kernel_variant": r64["variant"],
            "kernel": r64["kernel"],
            "tuning": r64["tuning"],
            "hbm_frac": r64["bytes_el"] * K / k64 / 1e9 / HBM_PEAK_GBS,
            "achieved_tflops": r64["flops_el"] * K / k64 / 1e12,
        }
        out["fp32_over_fp64"] = value / v64
        a, b = r32["state"], r64["state"]
        out["rel_l2_fp32_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        out["rel_l2_mass_fp32_vs_fp64"] = mass_norm(D, vx, E, fbc, order, a - b) / mass_norm(D, vx, E, fbc,
                                                                                              order, b)
        out["energy_fp64"] = r64["norms"]["energy"]
    out["energy_fp32"] = r32["norms"]["energy"]
    if not args.no_cpu:
        wall, csteps, threads = cpu_baseline(vx, E, fbc, order, q0, dt)
        wall1, csteps1, _ = cpu_baseline(vx, E, fbc, order, q0, dt, budget_s=8.0, threads=1)
        out["cpu_baseline"] = {
            "value": dof * 5 * csteps / wall,
            "unit": "DOF*RK-stages/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{csteps} LSERK4 steps (5 stages each) of the same {K}-tet p={order} mesh, "
                      f"FP64 CPU oracle, OpenMP {threads} threads, {wall:.2f} s",
            "value_1_thread": dof * 5 * csteps1 / wall1,
            "sample_1_thread": f"{csteps1} LSERK4 step(s) of the same mesh on 1 thread (the reference's "
                               f"serial Mult), {wall1:.2f} s",
        }
    if not args.headline_only:
        # the other configurations of the metric, measured in the same run
        out["configs"] = {"c1": bench_c1(D, args, cpu=not args.no_cpu), "c3": bench_c3(D, args),
                          "c4": bench_c4(D, args)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
