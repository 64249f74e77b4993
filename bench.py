"""Throughput of batched DPG local assembly on one MI355X.

One step = local assembly of every cell of the workload: K, rhs and the
static-condensation Schur complement (plus G for C1), inputs resident in HBM.

Default workload (the configuration the metric's target is quoted on):
C4, 3D Poisson ultraweak on 32^3 jittered hexes (+-5% h), p=3 (H1 order 4),
test order 5 (N=666, M=502), 16-mode sine load, FP64.

Prints one JSON line (see the field list in `main`). The CPU baseline is the
oracle (the FP64 C++ restatement of the reference algorithm, OpenMP over
cells) timed on this host on a bounded sample of the same cells; the same
sample gives the error of the GPU result against the FP64 oracle.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2.5e15, "hbm": 8.0e12}
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "fp16x3": 3, "int8split": 4}

CONFIGS = {
    "c4": dict(name="poisson3d_hex32_p3_test5", form=("poisson", 3), topo="hex", n=(32, 32, 32), jitter=0.05,
               k=4, kt=5, modes=16, seed=0xC0FFEE + 4),
    "c2": dict(name="stokes2d_quad128_p3_test5", form=("stokes", 2), topo="quad", n=(128, 128), jitter=0.10,
               k=4, kt=5, modes=8, seed=0xC0FFEE + 2),
    "c1": dict(name="poisson2d_quad16_p2_test4", form=("poisson", 2), topo="quad", n=(16, 16), jitter=0.10,
               k=3, kt=4, modes=8, seed=0xC0FFEE + 1),
    "c3": dict(name="cd2d_hp_graded_49296quads_128tris_p1-6_dp1-3", hp=True, n0=16, levels=10, tri_cols=4),
}


def build_inputs(cfg):
    from camellia2_amd import capi, dpg

    f = dpg.form(cfg["form"][0], cfg["form"][1])
    if cfg["topo"] == "hex":
        b = dpg.bucket(capi.HEX, cfg["k"], cfg["kt"])
        nodes, par = dpg.mesh_hex(*cfg["n"], cfg["jitter"], cfg["seed"])
        load = dpg.load_sine_modes(3, 1, cfg["modes"], 6, cfg["seed"] + 100)
    else:
        b = dpg.bucket(capi.QUAD, cfg["k"], cfg["kt"])
        nodes, par = dpg.mesh_quad(*cfg["n"], cfg["jitter"], cfg["seed"])
        ncomp = 2 if cfg["form"][0] == "stokes" else 1
        load = dpg.load_sine_modes(2, ncomp, cfg["modes"], 6, cfg["seed"] + 100)
    return f, b, nodes, par, load


ORACLE_BUILD = "g++ -O3 -march=x86-64-v3 -fopenmp, textbook loops (no BLAS/LAPACK on the host)"


def host_info():
    """nproc and the lscpu topology of the host the CPU baseline ran on."""
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keys = (("Model name", "model"), ("Socket(s)", "sockets"), ("Core(s) per socket", "cores_per_socket"),
            ("Thread(s) per core", "threads_per_core"))
    for line in out.splitlines():
        for key, name in keys:
            if line.startswith(key + ":"):
                info[name] = line.split(":", 1)[1].strip()
    return info


def cpu_baseline(f, b, nodes, par, load, gpu_K, target_s=2.5, reps=5, extra=None):
    """Oracle on a bounded sample of the workload's cells: the median of `reps`
    runs on all host threads (about target_s each) and one run on one thread
    (about target_s), plus the GPU-vs-oracle error on the sample (also for the
    leading cells' K of each entry of `extra`, name -> device tensor)."""
    from oracle import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ff = oracle.form_like(f)
    want = ("K", "rhs", "S", "g")
    # calibrate on one cell per thread, then size the sample to ~target_s
    n0 = max(1, threads)
    t0 = time.perf_counter()
    oracle.assemble(ff, b, nodes[:n0], par[:n0], None, load, want=want, nthreads=threads)
    per_cell = (time.perf_counter() - t0) / n0
    n = int(min(nodes.shape[0], max(n0, target_s / max(per_cell, 1e-9))))
    n = max(n0, (n // threads) * threads)
    times, o = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        o = oracle.assemble(ff, b, nodes[:n], par[:n], None, load, want=want, nthreads=threads)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    n1 = int(max(1, min(n, target_s / max(per_cell * threads, 1e-9))))
    t0 = time.perf_counter()
    oracle.assemble(ff, b, nodes[:n1], par[:n1], None, load, want=want, nthreads=1)
    dt1 = time.perf_counter() - t0
    def max_rel(Kdev):
        m = min(n, Kdev.shape[0])
        Kg = Kdev[:m].cpu().numpy()
        return float(np.max(np.linalg.norm((Kg - o["K"][:m]).reshape(m, -1), axis=1) /
                            np.linalg.norm(o["K"][:m].reshape(m, -1), axis=1))), m

    err, _ = max_rel(gpu_K)
    extra_err = {k: max_rel(v) for k, v in (extra or {}).items()}
    return {"value": n / dt, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"first {n} cells of the same mesh, oracle K+rhs+Schur, median of {reps} runs "
                      f"({dt:.2f} s wall each), {threads} OpenMP threads",
            "one_core_value": n1 / dt1, "one_core_sample": f"first {n1} cells, 1 thread, {dt1:.2f} s",
            "build": ORACLE_BUILD, "host": host_info()}, err, extra_err


def cpu_baseline_hp(f, groups, load, gpu_K, target_s=2.5, reps=5):
    """Oracle on a stratified sample of the hp mesh (the same share of every
    element type): the median of `reps` runs on all host threads (about
    target_s each) and one run on one thread; plus the GPU-vs-oracle error of K
    on that sample and, for each type's worst cell, the GPU's and the oracle's
    distance from K summed and solved in 80-bit (the oracle's is what FP64
    reaches on that cell)."""
    from oracle import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ff = oracle.form_like(f)
    want = ("K", "rhs", "S", "g")
    per = []
    for g in groups:
        n = min(threads, g["ncells"])
        t0 = time.perf_counter()
        oracle.assemble(ff, g["bucket"], g["nodes"][:n], g["parities"][:n], g["coefs"][:n], load,
                        want=want, nthreads=threads)
        per.append((time.perf_counter() - t0) / n)
    full = sum(t * g["ncells"] for t, g in zip(per, groups))

    def sample(frac):
        return [min(g["ncells"], max(1, int(round(frac * g["ncells"])))) for g in groups]

    def run(ns, nthreads):
        outs = []
        t0 = time.perf_counter()
        for g, n in zip(groups, ns):
            outs.append(oracle.assemble(ff, g["bucket"], g["nodes"][:n], g["parities"][:n], g["coefs"][:n], load,
                                        want=want, nthreads=nthreads))
        return time.perf_counter() - t0, outs

    frac = min(1.0, target_s / full)
    ns = sample(frac)
    times, outs = [], None
    for _ in range(reps):
        dt_, outs = run(ns, threads)
        times.append(dt_)
    dt = float(np.median(times))
    frac1 = min(1.0, target_s / (full * threads))
    ns1 = sample(frac1)
    dt1, _ = run(ns1, 1)
    err = 0.0
    worst = {"gpu_vs_80bit": 0.0, "oracle_vs_80bit": 0.0, "ratio_to_floor": 0.0}
    for bi, (g, n, o) in enumerate(zip(groups, ns, outs)):
        Kg = gpu_K[bi][:n].cpu().numpy()
        e = np.linalg.norm((Kg - o["K"]).reshape(n, -1), axis=1) / np.linalg.norm(o["K"].reshape(n, -1), axis=1)
        err = max(err, float(e.max()))
        c = int(np.argmax(e))
        o1, x1 = oracle.with_extended(ff, g["bucket"], g["nodes"][c:c + 1], g["parities"][c:c + 1],
                                      g["coefs"][c:c + 1], load)
        nx = np.linalg.norm(x1["K"][0])
        eg = float(np.linalg.norm(Kg[c] - x1["K"][0]) / nx)
        eo = float(np.linalg.norm(o1["K"][0] - x1["K"][0]) / nx)
        worst["gpu_vs_80bit"] = max(worst["gpu_vs_80bit"], eg)
        worst["oracle_vs_80bit"] = max(worst["oracle_vs_80bit"], eo)
        worst["ratio_to_floor"] = max(worst["ratio_to_floor"], eg / max(eo, 1e-16))
    n_tot, n1_tot = sum(ns), sum(ns1)
    return {"value": n_tot / dt, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"{n_tot} cells, the first {frac:.1%} of every element type, oracle K+rhs+Schur, "
                      f"median of {reps} runs ({dt:.2f} s wall each), {threads} OpenMP threads",
            "one_core_value": n1_tot / dt1,
            "one_core_sample": f"{n1_tot} cells, the first {frac1:.2%} of every element type, 1 thread, {dt1:.2f} s",
            "build": ORACLE_BUILD, "host": host_info()}, err, worst


def bench_hp(args, cfg):
    """hp mesh: every element type assembled by one captured graph launch,
    against per-bucket plain launches (default and autotuned tiles)."""
    import torch

    from camellia2_amd import capi, dpg
    from camellia2_amd.workloads import cd_hp, hp_groups

    f, cells, load = cd_hp(cfg["n0"], cfg["levels"], tri_cols=cfg["tri_cols"])
    groups = hp_groups(cells)
    buckets = [g["bucket"] for g in groups]
    C = sum(g["ncells"] for g in groups)
    ntri = sum(g["ncells"] for g in groups if g["topology"] == capi.TRI)
    dev = torch.device("cuda:0")
    ws = args.workspace_gib * 2 ** 30
    infos = [dpg.sizes(f, b) for b in buckets]
    flops = sum(g["ncells"] * i.flops_per_cell for g, i in zip(groups, infos))
    ins = [(torch.from_numpy(g["nodes"]).to(dev), torch.from_numpy(g["parities"]).to(dev),
            torch.from_numpy(g["coefs"]).to(dev)) for g in groups]

    def alloc():
        return [{"K": torch.empty((g["ncells"], i.M, i.M), dtype=torch.float64, device=dev),
                 "rhs": torch.empty((g["ncells"], i.M), dtype=torch.float64, device=dev),
                 "schur": torch.empty((g["ncells"], i.T, i.T), dtype=torch.float64, device=dev),
                 "schur_rhs": torch.empty((g["ncells"], i.T), dtype=torch.float64, device=dev)}
                for g, i in zip(groups, infos)]

    out_e, out_g = alloc(), alloc()

    def eager(asm):
        for bi, (n_, p_, c_) in enumerate(ins):
            asm.local_stiffness_and_rhs(bi, n_, p_, c_, load, want_schur=True, want_info=False, out=out_e[bi])

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    A = dpg.LocalAssembler(f, buckets, workspace_bytes=ws)
    dt_plain = timed(lambda: eager(A), args.steps)
    t0 = time.perf_counter()
    tiles = [A.autotune(bi, n_, p_, c_, load, sample=512) for bi, (n_, p_, c_) in enumerate(ins)]
    t_tune = time.perf_counter() - t0
    dt_tuned = timed(lambda: eager(A), args.steps)
    items = [dict(bucket=bi, nodes=n_, parities=p_, coefs=c_, load=load, want_schur=True, want_info=True,
                  out=out_g[bi]) for bi, (n_, p_, c_) in enumerate(ins)]
    G = dpg.AssemblyGraph(A, items, lanes=args.lanes)
    dt_graph = timed(G.launch, args.steps)
    same = all(torch.equal(out_g[bi][k], out_e[bi][k]) for bi in range(len(buckets))
               for k in ("K", "rhs", "schur", "schur_rhs"))
    nbad = int(sum(int((o["info"] != 0).sum()) for o in G.outputs()))
    gws = G.workspace_bytes()
    G.close()
    A.close()

    # per-stage kernel times of the tuned per-bucket launches (HIP events)
    Ap = dpg.LocalAssembler(f, buckets, workspace_bytes=ws, profile=True)
    for bi, t in enumerate(tiles):
        Ap.set_tiles(bi, *t)
    eager(Ap)
    torch.cuda.synchronize()
    tm, _ = Ap.timings(reset=True)
    Ap.close()
    stages = {k: {"ms": round(v["ms"], 3), "launches": v["launches"], "gflop": round(v["flops"] / 1e9, 3),
                  "gbyte": round(v["bytes"] / 1e9, 3)} for k, v in tm.items()}
    dom = max(["gram", "update", "panel"], key=lambda k: tm.get(k, {"ms": 0.0})["ms"])
    d = tm[dom]
    achieved = d["flops"] / (d["ms"] * 1e-3)
    roofline = {"bound": "mfma", "kernel": {"gram": "gemm_gram_f64", "update": "gemm_update_f64",
                                            "panel": "potrf_wave+panel_trsm"}[dom],
                "achieved": round(achieved / 1e12, 3), "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK["fp64"], 4), "traffic": None,
                "avg_launch_ms": round(d["ms"] / d["launches"], 4), "flop_per_launch": d["flops"] / d["launches"],
                "note": "summed over the element types' launches of this kernel (per-bucket plain launches)"
                        + ("; the panel stage runs on the FP64 vector units, whose published peak equals the "
                           "FP64 MFMA one" if dom == "panel" else "")}
    cps = C / dt_graph
    result = {
        "metric": "cells/s local DPG assembly, mixed-prec vs FP64 GPU form at ≤1e-10 rel err",
        "value": round(cps, 2),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_graph * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded graded hp mesh, per-cell p, dp, eps; sine-mode load)",
        "config": {"workload": cfg["name"], "cells": C, "quads": C - ntri, "triangles": ntri,
                   "element_types": len(buckets), "lanes": args.lanes,
                   "outputs": "K,rhs,schur,schur_rhs", "precision": "fp64",
                   "mode": "one hipGraph launch over all element types"},
        "gflops": round(flops / dt_graph / 1e9, 1),
        "algorithmic_flop_per_step": flops,
        "per_bucket_plain": {"cells_per_s_default_tiles": round(C / dt_plain, 2),
                             "cells_per_s_autotuned_tiles": round(C / dt_tuned, 2),
                             "graph_speedup_vs_default": round(dt_plain / dt_graph, 3),
                             "graph_speedup_vs_autotuned": round(dt_tuned / dt_graph, 3)},
        "graph_equals_plain_bitwise": bool(same),
        "non_spd_cells": nbad,
        "autotune": {"seconds": round(t_tune, 2),
                     "tiles": {("tri" if g["topology"] == capi.TRI else "quad") + f"_k{g['k']}_kt{g['kt']}": list(t)
                               for g, t in zip(groups, tiles)}},
        "graph_workspace_gib": round(gws / 2 ** 30, 2),
        "roofline": roofline,
        "stages": stages,
    }
    if not args.no_cpu:
        cpu, err, worst = cpu_baseline_hp(f, groups, load, [o["K"] for o in out_e])
        result["cpu_baseline"] = cpu
        result["rel_err_K_vs_fp64_oracle"] = err
        result["worst_cell_per_type_vs_80bit"] = worst
    attach_peaks(result, args.peaks)
    print(json.dumps(result))


def bench_solve_loop(args, f, b, nodes, cv, gdof, n, sysm, vals, rhs, solve, maxit=8000, tol=1e-10):
    """After the global assembly: CG on the condensed system with the trace
    fixed to 0 on the boundary (the sine loads vanish there), preconditioned by
    point Jacobi and by block Jacobi over the mesh entities (the solution used
    below), field recovery per cell and the energy error per cell, each timed."""
    import torch

    from camellia2_amd import dpg

    K, lrhs, par, load = solve
    fixed = torch.from_numpy(dpg.boundary_trace_mask(f, b, gdof, cv, n)).to("cuda:0")
    x = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters_j, rel_j = sysm.solve(vals, rhs, fixed, x, tol=tol, maxit=maxit)
    torch.cuda.synchronize()
    t_cg_j = time.perf_counter() - t0
    t0 = time.perf_counter()
    ent, nent = dpg.skeleton_entities(f, b, cv, n)
    sysm.set_blocks(ent, nent)
    t_blocks = time.perf_counter() - t0
    x.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters, rel = sysm.solve(vals, rhs, fixed, x, tol=tol, maxit=maxit)
    torch.cuda.synchronize()
    t_cg = time.perf_counter() - t0
    # one SpMV per iteration: values + columns + the row pointers, x gathered, y written
    spmv_bytes = 12.0 * sysm.nnz + 8.0 * (n + 1) + 16.0 * n
    s = dpg.sizes(f, b)
    C = nodes.shape[0]
    u = torch.zeros((C, s.M), dtype=torch.float64, device="cuda:0")
    u[:, s.F:] = x[torch.from_numpy(gdof.astype(np.int64)).to("cuda:0")]
    A = dpg.LocalAssembler(f, [b], workspace_bytes=args.workspace_gib * 2 ** 30)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nbad = A.recover_fields(0, K, lrhs, u)
    torch.cuda.synchronize()
    t_rec = time.perf_counter() - t0
    dn, dp = torch.from_numpy(nodes).to("cuda:0"), torch.from_numpy(par).to("cuda:0")
    t0 = time.perf_counter()
    eta2 = A.energy_error(0, dn, dp, u, None, load)
    torch.cuda.synchronize()
    t_eta = time.perf_counter() - t0
    A.close()
    flags, nmark = dpg.mark_cells(torch.sqrt(eta2).contiguous(), 0.2)
    return {"jacobi_cg_iterations": iters_j, "jacobi_cg_relres": rel_j, "jacobi_cg_s": round(t_cg_j, 3),
            "block_entities": nent, "block_setup_host_s": round(t_blocks, 3),
            "cg_iterations": iters, "cg_relres": rel, "cg_s": round(t_cg, 3),
            "cg_ms_per_iteration": round(t_cg / max(iters, 1) * 1e3, 3),
            "spmv_bytes": spmv_bytes, "spmv_gbs_upper": round(spmv_bytes / (t_cg / max(iters, 1)) / 1e9, 1),
            "recover_fields_s": round(t_rec, 3), "recover_non_spd": nbad, "energy_error_s": round(t_eta, 3),
            "energy_error_total": float(torch.sqrt(eta2.sum())), "cells_marked_at_0.2_max": nmark,
            "note": f"CG to relres {tol:g} (at most {maxit} iterations) preconditioned by block Jacobi over the "
                    "mesh entities (cg_*, including the block inverses) and by point Jacobi (jacobi_cg_*); the "
                    "per-iteration time includes three dot products read back to the host"}


def bench_global_assembly(args, f, b, nodes, schur, schur_rhs, cpu_target_s=2.0, solve=None):
    """The condensed global system of the same mesh: host numbering, device
    symbolic phase (CSR pattern + summation order), and the numeric assembly of
    every cell's Schur complement and load (timed, HBM roofline), with scipy's
    COO->CSR sum on a bounded sample of cells as the CPU stand-in."""
    import torch

    from camellia2_amd import dpg

    t0 = time.perf_counter()
    cv = dpg.vertex_ids(nodes)
    gdof, n = dpg.skeleton_numbering(f, b, cv)
    t_num = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sysm = dpg.CondensedSystem(gdof, n)
    torch.cuda.synchronize()
    t_sym = time.perf_counter() - t0
    C, T = gdof.shape
    vals, rhs = sysm.assemble(schur, schur_rhs)  # warm-up
    torch.cuda.synchronize()
    reps = max(1, args.steps)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        sysm.assemble(schur, schur_rhs)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    E = C * T * T
    # algorithmic bytes: Schur blocks and loads read once, their summation order
    # (4 B per contribution, 4 B per segment start), CSR values and load written
    nbytes = 8.0 * E + 4.0 * E + 4.0 * (sysm.nnz + 1) + 8.0 * sysm.nnz + 12.0 * C * T + 4.0 * (n + 1) + 8.0 * n
    gbs = nbytes / (ms * 1e-3) / 1e9
    res = {"global_dofs": n, "nnz": sysm.nnz, "contributions": E, "numbering_host_s": round(t_num, 3),
           "symbolic_device_s": round(t_sym, 3), "ms": round(ms, 3), "cells_per_s": round(C / (ms * 1e-3), 1),
           "roofline": {"bound": "hbm", "kernel": "ordered_segment_sum", "achieved": round(gbs, 1),
                        "peak": PEAK["hbm"] / 1e9, "unit": "GB/s", "frac": round(gbs * 1e9 / PEAK["hbm"], 4),
                        "bytes_per_launch": nbytes}}
    if args.peaks and "hbm_read_gbs" in args.peaks:
        res["roofline"]["frac_of_measured_copy"] = round(gbs / args.peaks["hbm_copy_gbs"], 4)
    if solve is not None:
        res["solve_loop"] = bench_solve_loop(args, f, b, nodes, cv, gdof, n, sysm, vals, rhs, solve)
    sysm.close()
    del vals, rhs
    if not args.no_cpu:
        import scipy.sparse as sp

        Sh = schur[:min(C, 64)].cpu().numpy()
        gd = gdof[:Sh.shape[0]]
        rows = np.repeat(gd, T, axis=1).ravel()
        cols = np.tile(gd, (1, T)).ravel()
        t0 = time.perf_counter()
        sp.coo_matrix((Sh.ravel(), (rows, cols)), shape=(n, n)).tocsr()
        per = (time.perf_counter() - t0) / Sh.shape[0]
        nc = int(min(C, max(Sh.shape[0], cpu_target_s / max(per, 1e-9))))
        Sh = schur[:nc].cpu().numpy()
        gd = gdof[:nc]
        rows = np.repeat(gd, T, axis=1).ravel()
        cols = np.tile(gd, (1, T)).ravel()
        t0 = time.perf_counter()
        sp.coo_matrix((Sh.ravel(), (rows, cols)), shape=(n, n)).tocsr()
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": nc / dt, "unit": "cells/s", "cores": 1, "kind": "port",
                               "sample": f"scipy coo_matrix(...).tocsr() of the first {nc} cells' Schur blocks, "
                                         f"{dt:.2f} s"}
    return res


def measured_peaks():
    """Sustained FP64 MFMA / FP64 FMA / HBM rates of this GPU from
    camellia2_amd/lib/dpg_peaks (a child process), or None if it is not built."""
    exe = os.path.join(ROOT, "camellia2_amd", "lib", "dpg_peaks")
    if not os.path.exists(exe):
        return None
    try:
        r = subprocess.run([exe], capture_output=True, text=True, timeout=180)
        if r.returncode != 0:
            return {"error": r.stderr.strip()[-300:]}
        return json.loads(r.stdout.strip().splitlines()[-1])
    except (OSError, subprocess.SubprocessError, ValueError, IndexError) as e:
        return {"error": str(e)}


def attach_peaks(result, peaks):
    """measured peaks, and the dominant kernel's fraction of the measured FP64 MFMA rate"""
    if peaks is None:
        return
    result["measured_peaks"] = peaks
    if "fp64_mfma_tflops" in peaks and peaks["fp64_mfma_tflops"] > 0:
        r = result["roofline"]
        r["frac_of_measured_peak"] = round(r["achieved"] / peaks["fp64_mfma_tflops"], 4)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="fp64", choices=sorted(PRECISIONS))
    ap.add_argument("--compare", default="int8split,fp32,bf16x3,fp16x3",
                    help="mixed precisions timed against the FP64 form (empty: none)")
    ap.add_argument("--workspace-gib", type=float, default=64.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-schur", action="store_true")
    ap.add_argument("--lanes", type=int, default=4, help="hp config: concurrent streams inside the graph")
    ap.add_argument("--tiles", default="",
                    help="GEMM tiles (64 or 32) for the Gram, B and Cholesky-update contractions, e.g. 64,32,64 "
                         "(single-type configurations; default: chosen by the plan autotuner)")
    ap.add_argument("--sumfact", action="store_true",
                    help="opt-in sum-factorised G (fewer flops, K accurate only to ~1e-12..1e-1; not the default)")
    ap.add_argument("--no-peaks", action="store_true", help="skip the peak-rate microbenchmark")
    ap.add_argument("--no-assembly", action="store_true",
                    help="skip the condensed global system (numbering, CSR pattern, assembly)")
    ap.add_argument("--calibrate-copy-gib", type=float, default=0.0,
                    help="also run a plain device copy of this size (counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (local assembly shards by "
              "cells with no exchange between GPUs; a multi-GPU run is one independent process per GPU)",
              file=sys.stderr)
        sys.exit(2)
    # the microbenchmark runs as a child process before this one touches the GPU
    args.peaks = None if args.no_peaks else measured_peaks()

    import torch

    from camellia2_amd import dpg

    cfg = CONFIGS[args.config]
    if cfg.get("hp"):
        return bench_hp(args, cfg)
    f, b, nodes, par, load = build_inputs(cfg)
    info = dpg.sizes(f, b)
    C = nodes.shape[0]
    dev = torch.device("cuda:0")
    d_nodes = torch.from_numpy(nodes).to(dev)
    d_par = torch.from_numpy(par).to(dev)
    ws = args.workspace_gib * 2 ** 30
    A = dpg.LocalAssembler(f, [b], workspace_bytes=ws, precision=PRECISIONS[args.precision])
    # GEMM tiles: given, or the plan's autotuner on the first cells (untimed set-up)
    t_tune = None
    if args.tiles:
        tiles = [int(t) for t in args.tiles.split(",")]
    elif args.sumfact:
        tiles = list(A.tiles(0)[:3])
    else:
        t0 = time.perf_counter()
        tiles = list(A.autotune(0, d_nodes, d_par, None, load, sample=min(C, 512))[:3])
        t_tune = round(time.perf_counter() - t0, 2)
    A.set_tiles(0, *tiles, 1 if args.sumfact else 0)
    out = {
        "K": torch.empty((C, info.M, info.M), dtype=torch.float64, device=dev),
        "rhs": torch.empty((C, info.M), dtype=torch.float64, device=dev),
    }
    want_schur = not args.no_schur
    if want_schur:
        out["schur"] = torch.empty((C, info.T, info.T), dtype=torch.float64, device=dev)
        out["schur_rhs"] = torch.empty((C, info.T), dtype=torch.float64, device=dev)

    def step(asm):
        return asm.local_stiffness_and_rhs(0, d_nodes, d_par, None, load, want_schur=want_schur, want_info=False,
                                           out=out)

    for _ in range(args.warmup):
        step(A)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(A)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    cells_per_s = C * args.steps / dt
    A.close()

    # one more step on a plan that times every launch with HIP events on the
    # launch stream (kept out of the timed steps above)
    Ap = dpg.LocalAssembler(f, [b], workspace_bytes=ws, profile=True, precision=PRECISIONS[args.precision])
    Ap.set_tiles(0, *tiles, 1 if args.sumfact else 0)
    step(Ap)
    torch.cuda.synchronize()
    tm, _ = Ap.timings(reset=True)
    Ap.close()
    stages = {k: {"ms": round(v["ms"], 3), "launches": v["launches"], "gflop": round(v["flops"] / 1e9, 3),
                  "gbyte": round(v["bytes"] / 1e9, 3)} for k, v in tm.items()}
    dom = max(["gram", "update"], key=lambda k: tm.get(k, {"ms": 0.0})["ms"])
    d = tm[dom]
    achieved = d["flops"] / (d["ms"] * 1e-3)
    roofline = {"bound": "mfma", "kernel": {"gram": "gemm_gram_f64", "update": "gemm_update_f64"}[dom],
                "achieved": round(achieved / 1e12, 3), "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK["fp64"], 4), "traffic": None,
                "avg_launch_ms": round(d["ms"] / d["launches"], 4),
                "flop_per_launch": d["flops"] / d["launches"]}
    traffic_file = os.path.join(ROOT, "profiles", f"traffic_{args.config}_{args.precision}.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as fh:
            tr = json.load(fh)
        k = tr.get(roofline["kernel"])
        roofline["traffic"] = None if k is None else k["bytes"]

    result = {
        "metric": "cells/s local DPG assembly, mixed-prec vs FP64 GPU form at ≤1e-10 rel err",
        "value": round(cells_per_s, 2),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else f"{args.precision} contractions, f64 factorisation",
        "data": "synthetic (seeded jittered mesh and sine-mode load)",
        "config": {"workload": cfg["name"], "cells": C, "N": info.N, "M": info.M, "T": info.T,
                   "outputs": "K,rhs" + (",schur,schur_rhs" if want_schur else ""),
                   "precision": args.precision, "gram": "sum-factorised (opt-in)" if args.sumfact else "dense",
                   "tiles": {"gram": tiles[0], "bmat": tiles[1], "update": tiles[2],
                             "chosen_by": "--tiles" if args.tiles else ("plan default" if args.sumfact else
                                                                         f"autotuner ({t_tune} s, untimed)")}},
        "gflops": round(info.flops_per_cell * cells_per_s / 1e9, 1),
        "algorithmic_flop_per_cell": info.flops_per_cell,
        "roofline": roofline,
        "stages": stages,
    }
    if want_schur and not args.no_assembly and args.precision == "fp64":
        solve = (out["K"], out["rhs"], par, load) if cfg["form"][0] == "poisson" else None
        result["global_assembly"] = bench_global_assembly(args, f, b, nodes, out["schur"], out["schur_rhs"],
                                                          solve=solve)
    # mixed-precision forms of the same step, timed the same way, with their
    # distance from the FP64 GPU result over all cells
    comps = [p for p in args.compare.split(",") if p and p != args.precision]
    if comps:
        ref_K = out["K"].clone() if args.precision == "fp64" else None
        mixed = {}
        heads = {}
        for p in comps:
            Am = dpg.LocalAssembler(f, [b], workspace_bytes=ws, precision=PRECISIONS[p])
            Am.set_tiles(0, *tiles, 0)  # the FP64 GEMMs of the mixed forms take the same tiles
            step(Am)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                step(Am)
            torch.cuda.synchronize()
            dtm = time.perf_counter() - t0
            Am.close()
            entry = {"cells_per_s": round(C * args.steps / dtm, 2), "ms_per_step": round(dtm / args.steps * 1e3, 3),
                     "speedup_vs_fp64": round(ms_per_step / (dtm / args.steps * 1e3), 3)}
            if ref_K is not None:
                rel = []
                for c0 in range(0, C, 1024):  # chunks: no full-size temporary
                    a, r = out["K"][c0:c0 + 1024], ref_K[c0:c0 + 1024]
                    rel.append(torch.linalg.matrix_norm(a - r) / torch.linalg.matrix_norm(r))
                rel = torch.cat(rel)
                fin = torch.isfinite(rel)
                entry["nonfinite_cells"] = int((~fin).sum())  # low-precision G not SPD / overflow
                rel = rel[fin]
                entry["rel_err_K_vs_fp64_gpu_max"] = float(rel.max()) if rel.numel() else None
                entry["rel_err_K_vs_fp64_gpu_median"] = float(rel.median()) if rel.numel() else None
            mixed[p] = entry
            heads[p] = out["K"][:min(C, 2048)].clone()  # for the distance from the oracle
        if ref_K is not None:
            out["K"].copy_(ref_K)
            del ref_K
        result["mixed_precision"] = mixed
    if not args.no_cpu:
        cpu, err, extra_err = cpu_baseline(f, b, nodes, par, load, out["K"], extra=heads if comps else None)
        result["cpu_baseline"] = cpu
        result["rel_err_K_vs_fp64_oracle"] = err
        for p, (e, m) in extra_err.items():
            result["mixed_precision"][p]["rel_err_K_vs_fp64_oracle_max"] = e
            result["mixed_precision"][p]["oracle_cells_compared"] = m
    if args.calibrate_copy_gib > 0:
        n = int(args.calibrate_copy_gib * 2 ** 30 // 8)
        src = torch.ones(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        result["calibration_copy_bytes"] = n * 8
        del src, dst
    attach_peaks(result, args.peaks)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
