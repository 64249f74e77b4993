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
_T_START = time.perf_counter()


def progress(msg):
    """One line on stderr when a phase of the run ends (the JSON result stays
    the only line on stdout)."""
    print(f"bench [{time.perf_counter() - _T_START:7.1f} s] {msg}", file=sys.stderr, flush=True)


PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "fp16x3": 3, "int8split": 4}

CONFIGS = {
    "c4": dict(name="poisson3d_hex32_p3_test5", form=("poisson", 3), topo="hex", n=(32, 32, 32), jitter=0.05,
               k=4, kt=5, modes=16, seed=0xC0FFEE + 4, cells_80bit=8),
    "c2": dict(name="stokes2d_quad128_p3_test5", form=("stokes", 2), topo="quad", n=(128, 128), jitter=0.10,
               k=4, kt=5, modes=8, seed=0xC0FFEE + 2, cells_80bit=64),
    "c1": dict(name="poisson2d_quad16_p2_test4", form=("poisson", 2), topo="quad", n=(16, 16), jitter=0.10,
               k=3, kt=4, modes=8, seed=0xC0FFEE + 1, cells_80bit=256),
    "c3": dict(name="cd2d_hp_graded_46215quads_6290tris_p1-6_dp1-3", hp=True, n0=16, levels=10, tri_cols=4,
               tri_rows=1),
    "c5": dict(name="cdr2d_quad64_p2_test4_adjoint", adjoint=True, n=(64, 64), jitter=0.10, k=3, kt=4,
               seed=0xC0FFEE + 5, cells_80bit=512),
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


def cpu_baseline(f, b, nodes, par, load, coefs=None, target_s=2.5, reps=5):
    """Oracle on a bounded sample of the workload's cells: the median of `reps`
    runs on all host threads (about target_s each) and one run on one thread
    (about target_s). Returns the baseline entry, the oracle outputs (K, rhs,
    S, g) of the sample and its size (the first n cells)."""
    from oracle import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ff = oracle.form_like(f)
    want = ("K", "rhs", "S", "g")
    cs = (lambda m: None) if coefs is None else (lambda m: coefs[:m])
    # calibrate on one cell per thread, then size the sample to ~target_s
    n0 = max(1, threads)
    t0 = time.perf_counter()
    oracle.assemble(ff, b, nodes[:n0], par[:n0], cs(n0), load, want=want, nthreads=threads)
    per_cell = (time.perf_counter() - t0) / n0
    n = int(min(nodes.shape[0], max(n0, target_s / max(per_cell, 1e-9))))
    n = max(n0, (n // threads) * threads)
    times, o = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        o = oracle.assemble(ff, b, nodes[:n], par[:n], cs(n), load, want=want, nthreads=threads)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    n1 = int(max(1, min(n, target_s / max(per_cell * threads, 1e-9))))
    t0 = time.perf_counter()
    oracle.assemble(ff, b, nodes[:n1], par[:n1], cs(n1), load, want=want, nthreads=1)
    dt1 = time.perf_counter() - t0
    return {"value": n / dt, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"first {n} cells of the same mesh, oracle K+rhs+Schur, median of {reps} runs "
                      f"({dt:.2f} s wall each), {threads} OpenMP threads",
            "one_core_value": n1 / dt1, "one_core_sample": f"first {n1} cells, 1 thread, {dt1:.2f} s",
            "build": ORACLE_BUILD, "host": host_info()}, o, n


def reference_sample(f, b, nodes, par, coefs, load, o, n, n_x, u=None):
    """What the GPU outputs are measured against on the leading cells: the FP64
    oracle's K, rhs, S, g (`o`, first n cells) and eta2 of the trial
    coefficients u; on the first n_x cells the 80-bit evaluation (G, B, l summed
    from the same FP64 point values, factored, solved and condensed in 80-bit)
    and the oracle's own distance from it, the FP64 floor."""
    from oracle import oracle

    ff = oracle.form_like(f)
    cs = (lambda m: None) if coefs is None else (lambda m: coefs[:m])
    ox, xx = oracle.with_extended(ff, b, nodes[:n_x], par[:n_x], cs(n_x), load, schur=True)
    ref = {"o": o, "n": n, "ox": ox, "x": xx, "n_x": n_x}
    if u is not None:
        ref["eta_o"] = oracle.energy_error(ff, b, nodes[:n], par[:n], u[:n], cs(n), load)
        ref["eta_x"] = oracle.eta2_extended(ff, b, nodes[:n_x], par[:n_x], u[:n_x], cs(n_x), load)
    return ref


ACC_KEYS = (("K", "K"), ("rhs", "rhs"), ("schur", "S"), ("schur_rhs", "g"))


def _max_rel(a, b):
    """max over cells (rows) of the relative distance; cells whose output is not
    finite are left out (None when none is finite)"""
    m = a.shape[0]
    a = a.reshape(m, -1)
    b = b.reshape(m, -1)
    e = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    e = e[np.isfinite(e)]
    return float(np.max(e)) if e.size else None


def accuracy(ref, gpu):
    """Max over the sample cells of the relative Frobenius distance of each GPU
    output (gpu: name -> leading cells as device tensors or arrays; eta2 by
    relative difference) from the FP64 oracle and from the 80-bit evaluation,
    next to the oracle's own distance from the 80-bit evaluation."""
    res = {"cells_vs_fp64_oracle": ref["n"], "cells_vs_80bit": ref["n_x"], "vs_fp64_oracle": {}, "vs_80bit": {},
           "oracle_vs_80bit": {}}
    host = {k: (v.cpu().numpy() if hasattr(v, "cpu") else np.asarray(v)) for k, v in gpu.items()}
    pairs = list(ACC_KEYS) + ([("eta2", None)] if "eta2" in host and "eta_o" in ref else [])
    for gk, ok in pairs:
        if gk not in host:
            continue
        g = host[gk]
        bad = int((~np.isfinite(g.reshape(g.shape[0], -1))).any(axis=1).sum())
        if bad:
            res.setdefault("nonfinite_cells", {})[gk] = bad
        o = ref["eta_o"] if gk == "eta2" else ref["o"][ok]
        x = ref["eta_x"] if gk == "eta2" else ref["x"][ok]
        ox = None if gk == "eta2" else ref["ox"][ok]
        m = min(ref["n"], g.shape[0])
        mx = min(ref["n_x"], g.shape[0])
        res["vs_fp64_oracle"][gk] = _max_rel(g[:m], o[:m])
        res["vs_80bit"][gk] = _max_rel(g[:mx], x[:mx])
        res["oracle_vs_80bit"][gk] = _max_rel((ref["eta_o"][:mx] if gk == "eta2" else ox[:mx]), x[:mx])
    return res


def cpu_baseline_hp(f, groups, load, gpu_K, target_s=2.5, reps=5):
    """Oracle on a stratified sample of the hp mesh (the same share of every
    element type): the median of `reps` runs on all host threads (about
    target_s each) and one run on one thread; plus the GPU-vs-oracle error of K
    on that sample and, for each type's worst cell, the GPU's and the oracle's
    distance from K summed and solved in 80-bit (the oracle's, the largest over
    the cell and three copies moved at the rounding level of its geometry, is
    what FP64 reaches on that cell)."""
    from camellia2_amd import capi
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
        # the floor as the tests take it: the largest distance from the 80-bit K
        # of the oracle's FP64 K on the cell and on three copies moved at the
        # rounding level of its geometry (the GPU rounds its point values
        # independently); triangles with the oracle's test basis of the
        # library's family
        nd = g["nodes"][c:c + 1]
        diam = np.max(np.linalg.norm(nd - nd[:, :1], axis=2), axis=1)
        rng = np.random.default_rng(7 + bi)
        copies = [nd + 2.0 ** -51 * diam[:, None, None] * rng.uniform(-1.0, 1.0, nd.shape) for _ in range(3)]
        with oracle.tri_test_basis(1 if g["bucket"].topology == capi.TRI else 0):
            o1, x1 = oracle.with_extended(ff, g["bucket"], nd, g["parities"][c:c + 1], g["coefs"][c:c + 1], load)
            moved = [oracle.assemble(ff, g["bucket"], m, g["parities"][c:c + 1], g["coefs"][c:c + 1], load,
                                     want=("K",))["K"][0] for m in copies]
        nx = np.linalg.norm(x1["K"][0])
        eg = float(np.linalg.norm(Kg[c] - x1["K"][0]) / nx)
        eo = max(float(np.linalg.norm(k - x1["K"][0]) / nx) for k in [o1["K"][0]] + moved)
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

    f, cells, load = cd_hp(cfg["n0"], cfg["levels"], tri_cols=cfg["tri_cols"], tri_rows=cfg["tri_rows"])
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
    progress("per-type launches (default tiles) done")
    t0 = time.perf_counter()
    tiles = [A.autotune(bi, n_, p_, c_, load, sample=512) for bi, (n_, p_, c_) in enumerate(ins)]
    t_tune = time.perf_counter() - t0
    dt_tuned = timed(lambda: eager(A), args.steps)
    progress("autotune and per-type launches (tuned tiles) done")
    items = [dict(bucket=bi, nodes=n_, parities=p_, coefs=c_, load=load, want_schur=True, want_info=True,
                  out=out_g[bi]) for bi, (n_, p_, c_) in enumerate(ins)]
    G = dpg.AssemblyGraph(A, items, lanes=args.lanes)
    dt_graph = timed(G.launch, args.steps)
    progress("graph launches done")
    same = all(torch.equal(out_g[bi][k], out_e[bi][k]) for bi in range(len(buckets))
               for k in ("K", "rhs", "schur", "schur_rhs"))
    nbad = int(sum(int((o["info"] != 0).sum()) for o in G.outputs()))
    gws = G.workspace_bytes()
    G.close()
    A.close()
    hp_system = None
    if not args.no_assembly:
        hp_system = bench_hp_system(args, f, groups, buckets, out_e, load)
        progress("hp condensed system done")

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
    dom = max(["gram", "update", "panel", "solve"], key=lambda k: tm.get(k, {"ms": 0.0})["ms"])
    d = tm[dom]
    achieved = d["flops"] / (d["ms"] * 1e-3)
    roofline = {"bound": "mfma", "kernel": {"gram": "gemm_gram_f64", "update": "gemm_update_f64",
                                            "panel": "potrf_wave+panel_trsm", "solve": "update_solve_f64"}[dom],
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
        "global_assembly": hp_system,
        "roofline": roofline,
        "stages": stages,
    }
    if not args.no_cpu:
        cpu, err, worst = cpu_baseline_hp(f, groups, load, [o["K"] for o in out_e])
        progress("CPU baseline done")
        result["cpu_baseline"] = cpu
        result["rel_err_K_vs_fp64_oracle"] = err
        result["worst_cell_per_type_vs_80bit"] = worst
    attach_peaks(result, args.peaks)
    print(json.dumps(result))


def c5_inputs(cfg):
    """2D CDR ultraweak on 64x64 jittered quads with per-cell coefficients:
    eps log-uniform on [1e-3, 1], beta = (cos t, sin t) * U[0.5, 2] with one
    seeded angle t, reaction U[0, 1]; the 8-mode sine load of the 16x16 Poisson
    workload; trial coefficients u ~ N(0, 1) per cell (a frozen input of eta)."""
    from camellia2_amd import capi, dpg

    seed = cfg["seed"]
    f = dpg.form("cdr", 2)
    b = dpg.bucket(capi.QUAD, cfg["k"], cfg["kt"])
    nodes, par = dpg.mesh_quad(*cfg["n"], cfg["jitter"], seed)
    C = nodes.shape[0]
    eps = 10.0 ** dpg.rng_uniform(seed + 1, -3.0, 0.0, C)
    theta = dpg.rng_uniform(seed + 2, 0.0, 2 * np.pi, 1)[0]
    mag = dpg.rng_uniform(seed + 3, 0.5, 2.0, C)
    react = dpg.rng_uniform(seed + 4, 0.0, 1.0, C)
    coefs = np.stack([eps, mag * np.cos(theta), mag * np.sin(theta), react], axis=1)
    load = dpg.load_sine_modes(2, 1, 8, 6, 0xC0FFEE + 1 + 100)
    M = dpg.sizes(f, b).M
    u = dpg.rng_normal(seed + 50, C * M).reshape(C, M)
    return f, b, nodes, par, coefs, load, u


def bench_c5(args, cfg):
    """Differentiation workload: one step = the forward local assembly (K, rhs,
    Schur complement, the factorisation kept) of every cell, then from the kept
    factorisation the energy error of u and the adjoint gradient of the summed
    indicator with respect to all four coefficients of every cell."""
    import torch

    from camellia2_amd import dpg

    f, b, nodes, par, coefs, load, u = c5_inputs(cfg)
    info = dpg.sizes(f, b)
    C = nodes.shape[0]
    dev = torch.device("cuda:0")
    dn, dp, dc, du = (torch.from_numpy(a).to(dev) for a in (nodes, par, coefs, u))
    ws = args.workspace_gib * 2 ** 30
    A = dpg.LocalAssembler(f, [b], workspace_bytes=ws)
    t0 = time.perf_counter()
    tiles = list(A.autotune(0, dn, dp, dc, load, sample=min(C, 512))[:3])
    t_tune = round(time.perf_counter() - t0, 2)
    out = {"K": torch.empty((C, info.M, info.M), dtype=torch.float64, device=dev),
           "rhs": torch.empty((C, info.M), dtype=torch.float64, device=dev),
           "schur": torch.empty((C, info.T, info.T), dtype=torch.float64, device=dev),
           "schur_rhs": torch.empty((C, info.T), dtype=torch.float64, device=dev)}
    res = {}

    def step(asm):
        asm.local_stiffness_and_rhs(0, dn, dp, dc, load, want_schur=True, want_info=False, out=out, want_factor=True)
        res["eta2"], res["grad"] = asm.adjoint(0, dn, dp, dc, load, u=du, factor=out["factor"],
                                               factor_scale=out["factor_scale"], want_eta2=True)

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    dt = timed(lambda: step(A), args.steps)
    progress(f"timed steps done: {dt * 1e3:.2f} ms per step")
    # the parts of the step on their own
    dt_fwd = timed(lambda: A.local_stiffness_and_rhs(0, dn, dp, dc, load, want_schur=True, want_info=False, out=out,
                                                      want_factor=True), args.steps)
    dt_adj = timed(lambda: A.adjoint(0, dn, dp, dc, load, u=du, factor=out["factor"],
                                     factor_scale=out["factor_scale"], want_eta2=True), args.steps)
    dt_eta_re = timed(lambda: A.energy_error(0, dn, dp, du, dc, load), args.steps)
    dt_adj_re = timed(lambda: A.energy_error_and_gradient(0, dn, dp, du, dc, load), args.steps)
    progress("parts timed")
    step(A)
    torch.cuda.synchronize()
    eta2, grad = res["eta2"].clone(), res["grad"].clone()
    # adjoint against central differences of the GPU forward (step 1e-6, every cell at once)
    fd_err = {}
    for k, name in enumerate(("eps", "beta_x", "beta_y", "reaction")):
        cp, cm = coefs.copy(), coefs.copy()
        cp[:, k] += 1e-6
        cm[:, k] -= 1e-6
        ep = A.energy_error(0, dn, dp, du, torch.from_numpy(cp).to(dev), load)
        em = A.energy_error(0, dn, dp, du, torch.from_numpy(cm).to(dev), load)
        fd = ((ep - em) / 2e-6).cpu().numpy()
        g = grad[:, k].cpu().numpy()
        fd_err[name] = float(np.linalg.norm(g - fd) / np.linalg.norm(fd))
    A.close()
    # per-stage kernel times of one step (HIP events on the launch stream)
    Ap = dpg.LocalAssembler(f, [b], workspace_bytes=ws, profile=True)
    Ap.set_tiles(0, *tiles)
    step(Ap)
    torch.cuda.synchronize()
    tm, _ = Ap.timings(reset=True)
    Ap.close()
    progress("profiled step done")
    stages = {k: {"ms": round(v["ms"], 4), "launches": v["launches"], "gflop": round(v["flops"] / 1e9, 4),
                  "gbyte": round(v["bytes"] / 1e9, 4)} for k, v in tm.items()}
    dom = max(tm, key=lambda k: tm[k]["ms"])
    d = tm[dom]
    mfma = dom in ("gram", "update", "bmat", "solve")
    if mfma or d["flops"] / max(d["bytes"], 1.0) > PEAK["fp64"] / PEAK["hbm"]:
        achieved, peak, unit, bound = d["flops"] / (d["ms"] * 1e-3), PEAK["fp64"], "TFLOP/s", "mfma"
    else:
        achieved, peak, unit, bound = d["bytes"] / (d["ms"] * 1e-3), PEAK["hbm"], "GB/s", "hbm"
    scale = 1e12 if unit == "TFLOP/s" else 1e9
    kernels = {"gram": "gemm_gram_f64", "update": "gemm_update_f64", "bmat": "gemm_bmat_f64",
               "solve": "update_solve_f64",
               "panel": "potrf_wave+panel_trsm", "adjoint": "factor_backsolve+adjoint_kernel",
               "operand": "ip_operand_kernel+operand_kernel", "point": "point_kernel",
               "extract": "extract_sym+extract_factor", "scale": "apply_scaling"}
    roofline = {"bound": bound, "kernel": kernels.get(dom, dom), "achieved": round(achieved / scale, 4),
                "peak": peak / scale, "unit": unit, "frac": round(achieved / peak, 5), "traffic": None,
                "avg_launch_ms": round(d["ms"] / d["launches"], 5), "stage": dom,
                "note": "dominant stage of one profiled step; latency-bound at 4,096 small cells"}
    fwd_flops = info.flops_per_cell * C
    result = {
        "metric": "cells/s local DPG assembly, mixed-prec vs FP64 GPU form at ≤1e-10 rel err",
        "value": round(C / dt, 2),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded jittered mesh, per-cell eps, beta, reaction; sine-mode load; u ~ N(0,1))",
        "config": {"workload": cfg["name"], "cells": C, "N": info.N, "M": info.M, "T": info.T,
                   "outputs": "K,rhs,schur,schur_rhs,factor,eta2,grad(eta2 wrt 4 coefficients per cell)",
                   "precision": "fp64", "tiles": {"gram": tiles[0], "bmat": tiles[1], "update": tiles[2],
                                                  "chosen_by": f"autotuner ({t_tune} s, untimed)"}},
        "gflops_forward": round(fwd_flops / dt / 1e9, 2),
        "algorithmic_flop_per_cell_forward": info.flops_per_cell,
        "parts_ms": {"forward_with_factor": round(dt_fwd * 1e3, 4), "eta2_and_adjoint_from_kept_factor":
                     round(dt_adj * 1e3, 4), "eta2_reassembled": round(dt_eta_re * 1e3, 4),
                     "eta2_and_adjoint_reassembled": round(dt_adj_re * 1e3, 4)},
        "adjoint_vs_central_differences_h1e-6": fd_err,
        "roofline": roofline,
        "stages": stages,
    }
    if not args.no_cpu:
        from oracle import oracle

        cpu, o, n = cpu_baseline(f, b, nodes, par, load, coefs=coefs)
        t0 = time.perf_counter()
        oracle.energy_error(oracle.form_like(f), b, nodes[:n], par[:n], u[:n], coefs[:n], load)
        t_eta = time.perf_counter() - t0
        cpu["value_with_energy_error"] = n / (n / cpu["value"] + t_eta)
        cpu["note"] = ("the reference has no adjoint: the CPU baseline is the forward (K, rhs, Schur) plus, in "
                       "value_with_energy_error, the energy error of the same cells")
        progress("CPU baseline done")
        result["cpu_baseline"] = cpu
        ref = reference_sample(f, b, nodes, par, coefs, load, o, n, min(n, cfg["cells_80bit"]), u=u)
        gpu = {k: out[k][:n] for k in ("K", "rhs", "schur", "schur_rhs")}
        gpu["eta2"] = eta2[:n]
        result["accuracy"] = accuracy(ref, gpu)
        progress("accuracy against the oracle and the 80-bit evaluation done")
    attach_peaks(result, args.peaks)
    print(json.dumps(result))


def bench_hp_system(args, f, groups, buckets, out, load, maxit=8000, tol=1e-10):
    """The condensed system of the hp mesh (hanging sides, order jumps,
    triangles): constrained numbering (host), assembly of every element type's
    Schur blocks, block-Jacobi CG with the trace fixed to 0 on the boundary,
    the local trace/flux coefficients and field recovery, each timed."""
    import torch

    from camellia2_amd import dpg

    pts = np.concatenate([g["nodes"].reshape(-1, 2) for g in groups])
    X, inv = np.unique(pts, axis=0, return_inverse=True)
    C = sum(g["ncells"] for g in groups)
    cv = np.full((C, 4), -1, dtype=np.int32)
    cb = np.zeros(C, dtype=np.int32)
    pos = row = 0
    for b, g in enumerate(groups):
        n, V = g["nodes"].shape[0], g["nodes"].shape[1]
        cv[row:row + n, :V] = inv[pos:pos + n * V].reshape(n, V)
        cb[row:row + n] = b
        pos += n * V
        row += n
    t0 = time.perf_counter()
    hs = dpg.HpSystem(f, buckets, cb, cv, X)
    t_setup = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vals, rhs = hs.assemble([o["schur"] for o in out], [o["schur_rhs"] for o in out])
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    fixed, _, _, _ = hs.dofs()
    x = torch.zeros(hs.nglobal, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters, rel = hs.solve(vals, rhs, torch.from_numpy(fixed).to("cuda:0"), x, tol=tol, maxit=maxit)
    torch.cuda.synchronize()
    t_cg = time.perf_counter() - t0
    u = [torch.zeros((g["ncells"], dpg.sizes(f, bk).M), dtype=torch.float64, device="cuda:0")
         for g, bk in zip(groups, buckets)]
    A = dpg.LocalAssembler(f, buckets, workspace_bytes=args.workspace_gib * 2 ** 30)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hs.local(x, u)
    nbad = sum(A.recover_fields(b, out[b]["K"], out[b]["rhs"], u[b]) for b in range(len(buckets)))
    torch.cuda.synchronize()
    t_rec = time.perf_counter() - t0
    A.close()
    res = {"cells": C, "global_dofs": hs.nglobal, "nnz": hs.nnz, "entities": hs.nentities,
           "setup_host_s": round(t_setup, 3), "assembly_ms": round(t_asm * 1e3, 3), "cg_iterations": iters,
           "cg_relres": rel, "cg_s": round(t_cg, 3), "local_and_recovery_ms": round(t_rec * 1e3, 3),
           "recover_non_spd": nbad,
           "note": "minimum rule over hanging sides and order jumps (dpg_hp_system); CG with block Jacobi over "
                   f"vertices and sides to relres {tol:g} (at most {maxit} iterations), trace 0 on the boundary"}
    hs.close()
    return res


def bench_solve_loop(args, f, b, nodes, cv, gdof, n, sysm, vals, rhs, solve, schur, maxit=8000, tol=1e-10):
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
    # two levels: the same smoother and a correction on the lowest-order trace/flux level
    t0 = time.perf_counter()
    co = dpg.CoarseLevel(f, b, cv)
    t_coarse_setup = time.perf_counter() - t0
    x2 = torch.zeros_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters2, rel2 = sysm.solve_twolevel(co, schur, vals, rhs, fixed, x2, tol=tol, maxit=maxit, coarse_iters=20)
    torch.cuda.synchronize()
    t_2l = time.perf_counter() - t0
    diff_2l = float((x2 - x).abs().max() / x.abs().max())
    coarse_n, coarse_nnz = co.nglobal, co.nnz
    co.close()
    del x2
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
    # the energy error from a factorisation kept by the assembly (no re-assembly),
    # on the first cells (the kept factor of the whole mesh would not fit beside K)
    ch = min(C, 4096)
    fo, _ = A.local_stiffness_and_rhs(0, dn[:ch], dp[:ch], None, load, want_factor=True, want_info=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eta2_f = A.energy_error_factored(0, fo["factor"], u[:ch].contiguous())
    torch.cuda.synchronize()
    t_eta_f = time.perf_counter() - t0
    t0 = time.perf_counter()
    A.energy_error(0, dn[:ch], dp[:ch], u[:ch].contiguous(), None, load)
    torch.cuda.synchronize()
    t_eta_ch = time.perf_counter() - t0
    eta_f_diff = float(((eta2_f - eta2[:ch]).abs() / eta2[:ch].abs().clamp_min(1e-300)).max())
    del fo
    A.close()
    flags, nmark = dpg.mark_cells(torch.sqrt(eta2).contiguous(), 0.2)
    return {"jacobi_cg_iterations": iters_j, "jacobi_cg_relres": rel_j, "jacobi_cg_s": round(t_cg_j, 3),
            "block_entities": nent, "block_setup_host_s": round(t_blocks, 3),
            "cg_iterations": iters, "cg_relres": rel, "cg_s": round(t_cg, 3),
            "cg_ms_per_iteration": round(t_cg / max(iters, 1) * 1e3, 3),
            "twolevel": {"iterations": iters2, "relres": rel2, "s": round(t_2l, 3),
                         "ms_per_iteration": round(t_2l / max(iters2, 1) * 1e3, 3),
                         "coarse_dofs": coarse_n, "coarse_nnz": coarse_nnz, "coarse_cg_steps": 20,
                         "setup_host_s": round(t_coarse_setup, 3), "max_rel_diff_vs_block_jacobi_cg": diff_2l},
            "spmv_bytes": spmv_bytes, "spmv_gbs_upper": round(spmv_bytes / (t_cg / max(iters, 1)) / 1e9, 1),
            "recover_fields_s": round(t_rec, 3), "recover_non_spd": nbad, "energy_error_s": round(t_eta, 3),
            "energy_error_from_kept_factor": {"cells": ch, "s": round(t_eta_f, 5), "reassembled_s": round(t_eta_ch, 4),
                                              "max_rel_diff": eta_f_diff},
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
        res["solve_loop"] = bench_solve_loop(args, f, b, nodes, cv, gdof, n, sysm, vals, rhs, solve, schur)
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


def library_dgemm_tflops(n=8192, reps=10):
    """FP64 GEMM rate of the vendor BLAS through torch.matmul (n x n x n, the
    median of `reps` launches): a second ceiling for the FP64 MFMA kernels next
    to the register-only microbenchmark."""
    import torch

    a = torch.randn((n, n), dtype=torch.float64, device="cuda:0")
    b = torch.randn((n, n), dtype=torch.float64, device="cuda:0")
    c = torch.matmul(a, b)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        torch.matmul(a, b, out=c)
        e1.record()
        torch.cuda.synchronize()
        ts.append(e0.elapsed_time(e1) * 1e-3)
    del a, b, c
    return round(2.0 * n ** 3 / float(np.median(ts)) / 1e12, 2)


def attach_peaks(result, peaks):
    """measured peaks, and the dominant kernel's fraction of the highest measured
    FP64 MFMA rate (the library DGEMM on 8192^2 sustains more than a register-only
    MFMA loop does, so it is the ceiling when it was measured)"""
    if peaks is None:
        return
    result["measured_peaks"] = peaks
    r = result["roofline"]
    if r["unit"] == "GB/s":
        if peaks.get("hbm_copy_gbs", 0) > 0:
            r["frac_of_measured_peak"] = round(r["achieved"] / peaks["hbm_copy_gbs"], 4)
    else:
        ceiling = max(peaks.get("fp64_mfma_tflops", 0), peaks.get("library_dgemm_tflops", 0) or 0)
        if ceiling > 0:
            r["frac_of_measured_peak"] = round(r["achieved"] / ceiling, 4)
            r["measured_peak_tflops"] = ceiling


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
                    help="GEMM tiles (64 or 32; 128 for the updates) for the Gram, B and Cholesky-update contractions, e.g. 64,32,128 "
                         "(single-type configurations; default: chosen by the plan autotuner)")
    ap.add_argument("--gram", default="auto", choices=["auto", "dense", "sumfact"],
                    help="G from the dense MFMA Gram or the sum-factorised kernel (tensor cells); auto: the plan "
                         "autotuner's choice, which takes the sum-factorised G only where it is faster and moves K "
                         "no more than a second rounding of the dense G does")
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
    progress("peak-rate microbenchmark done")

    import torch

    from camellia2_amd import dpg

    if args.peaks is not None and "error" not in args.peaks:
        args.peaks["library_dgemm_tflops"] = library_dgemm_tflops()
        progress(f"library FP64 GEMM: {args.peaks['library_dgemm_tflops']} TFLOP/s")
    cfg = CONFIGS[args.config]
    if cfg.get("hp"):
        return bench_hp(args, cfg)
    if cfg.get("adjoint"):
        return bench_c5(args, cfg)
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
    sf = 0
    if args.gram == "auto" or not args.tiles:
        t0 = time.perf_counter()
        tuned = A.autotune(0, d_nodes, d_par, None, load, sample=min(C, 512))
        t_tune = round(time.perf_counter() - t0, 2)
        tiles, sf = list(tuned[:3]), int(tuned[3])
        progress(f"autotuner chose tiles {tiles}, sum-factorised G {sf}")
    if args.tiles:
        tiles = [int(t) for t in args.tiles.split(",")]
    if args.gram != "auto":
        sf = 1 if args.gram == "sumfact" else 0
    A.set_tiles(0, *tiles, sf)
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
    progress(f"timed steps done: {ms_per_step:.1f} ms per step")

    # one more step on a plan that times every launch with HIP events on the
    # launch stream (kept out of the timed steps above)
    Ap = dpg.LocalAssembler(f, [b], workspace_bytes=ws, profile=True, precision=PRECISIONS[args.precision])
    Ap.set_tiles(0, *tiles, sf)
    step(Ap)
    torch.cuda.synchronize()
    tm, _ = Ap.timings(reset=True)
    Ap.close()
    progress("profiled step done")
    stages = {k: {"ms": round(v["ms"], 3), "launches": v["launches"], "gflop": round(v["flops"] / 1e9, 3),
                  "gbyte": round(v["bytes"] / 1e9, 3)} for k, v in tm.items()}
    dom = max(["gram", "update", "solve"], key=lambda k: tm.get(k, {"ms": 0.0})["ms"])
    d = tm[dom]
    achieved = d["flops"] / (d["ms"] * 1e-3)
    roofline = {"bound": "mfma",
                "kernel": {"gram": "gemm_gram_f64", "update": "gemm_update_f64", "solve": "update_solve_f64"}[dom],
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
                   "precision": args.precision,
                   "gram": ("sum-factorised" if sf else "dense") + (" (autotuner)" if args.gram == "auto" else ""),
                   "tiles": {"gram": tiles[0], "bmat": tiles[1], "update": tiles[2],
                             "chosen_by": "--tiles" if args.tiles else f"autotuner ({t_tune} s, untimed)"}},
        "gflops": round(info.flops_per_cell * cells_per_s / 1e9, 1),
        "gflops_note": "the dense algorithm's flop count per cell times cells/s: a rate of algorithmic work, not a "
                       "utilisation (the Gram is computed as a lower triangle, and the sum-factorised G and B do far "
                       "fewer flops than the contractions counted); utilisation is the roofline's",
        "algorithmic_flop_per_cell": info.flops_per_cell,
        "roofline": roofline,
        "stages": stages,
    }
    if want_schur and not args.no_assembly and args.precision == "fp64":
        solve = (out["K"], out["rhs"], par, load) if cfg["form"][0] == "poisson" else None
        result["global_assembly"] = bench_global_assembly(args, f, b, nodes, out["schur"], out["schur_rhs"],
                                                          solve=solve)
        progress("condensed global system done")
    # the leading cells of every form's outputs, with the energy error of seeded
    # trial coefficients there, are kept for the distance from the oracle
    head = min(C, 1024 if cfg["topo"] == "hex" else 4096)
    u_head = dpg.rng_normal(cfg["seed"] + 55, head * info.M).reshape(head, info.M)
    d_u = torch.from_numpy(u_head).to(dev)

    def heads_of(asm):
        h = {k: out[k][:head].clone() for k in ("K", "rhs", "schur", "schur_rhs") if k in out}
        h["eta2"] = asm.energy_error(0, d_nodes[:head], d_par[:head], d_u, None, load)
        return h

    Ah = dpg.LocalAssembler(f, [b], workspace_bytes=ws, precision=PRECISIONS[args.precision])
    Ah.set_tiles(0, *tiles, sf)
    heads = {args.precision: heads_of(Ah)}
    Ah.close()
    # mixed-precision forms of the same step, timed the same way, with their
    # distance from the FP64 GPU result over all cells
    comps = [p for p in args.compare.split(",") if p and p != args.precision]
    if comps:
        ref_K = out["K"].clone() if args.precision == "fp64" else None
        mixed = {}
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
            heads[p] = heads_of(Am)
            Am.close()
            mixed[p] = entry
            progress(f"{p} form done: {entry['ms_per_step']} ms per step")
        if ref_K is not None:
            out["K"].copy_(ref_K)
            del ref_K
        result["mixed_precision"] = mixed
    if not args.no_cpu:
        cpu, o, n = cpu_baseline(f, b, nodes, par, load)
        progress("CPU baseline done")
        result["cpu_baseline"] = cpu
        result["rel_err_K_vs_fp64_oracle"] = _max_rel(out["K"][:n].cpu().numpy(), o["K"])
        ref = reference_sample(f, b, nodes, par, None, load, o, min(n, head), min(n, head, cfg["cells_80bit"]),
                               u=u_head)
        result["accuracy"] = accuracy(ref, heads[args.precision])
        for p in comps:
            result["mixed_precision"][p]["accuracy"] = accuracy(ref, heads[p])
        progress("accuracy against the oracle and the 80-bit evaluation done")
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
