"""hp meshes on the GPU: every (topology, p, dp) element type of the graded
convection-diffusion mesh against the oracle, triangles, GEMM tile choices, and
the all-bucket hipGraph launch against per-bucket plain launches (bit for bit)."""
import numpy as np
import pytest

from camellia2_amd import capi, dpg
from oracle import oracle

from helpers import cd_hp, hp_groups, linear_trial_coefficients, rel_fro, split_quads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def c3_case(n0, levels, tri_cols=1):
    f, cells, load = cd_hp(n0, levels, tri_cols=tri_cols, tri_rows=0)
    return f, hp_groups(cells), load


def rounding_copies(nodes, n, seed):
    """n copies of the cells with every vertex moved by up to 2^-51 times the
    cell diameter: inputs that differ at the rounding level of the geometry"""
    diam = np.max(np.linalg.norm(nodes - nodes[:, :1], axis=2), axis=1)
    rng = np.random.default_rng(seed)
    return [nodes + 2.0 ** -51 * diam[:, None, None] * rng.uniform(-1.0, 1.0, nodes.shape) for _ in range(n)]


def check_types_against_oracle(f, groups, load, want_schur=True, copies=3):
    """each element type vs the oracle: K, rhs (and the Schur complement) within
    max(1e-12, 4 x floor) of the 80-bit evaluation (G, B, l summed and solved in
    80-bit from the oracle's FP64 point values). floor = the largest distance
    from it of the oracle's FP64 outputs on the cells and on `copies` copies
    moved at the rounding level of their geometry: the GPU rounds its point
    values (Jacobians, weights) independently of the oracle, and on ill
    conditioned cells that difference alone moves K by the same order."""
    A = dpg.LocalAssembler(f, [g["bucket"] for g in groups])
    worst = 0.0
    for bi, g in enumerate(groups):
        print(f"type topo={g['topology']} k={g['k']} kt={g['kt']} cells={g['ncells']}", flush=True)
        out, nbad = A.local_stiffness_and_rhs(bi, dev(g["nodes"]), dev(g["parities"]), dev(g["coefs"]), load,
                                              want_schur=want_schur)
        torch.cuda.synchronize()
        res = {key: v.cpu().numpy() for key, v in out.items()}
        assert nbad == 0 and np.all(res["info"] == 0)
        tri = g["topology"] == capi.TRI
        args = (oracle.form_like(f), g["bucket"], g["nodes"], g["parities"], g["coefs"], load)
        # triangles: the oracle's test basis of the library's family (orthonormal,
        # centroid RT complement), so that both round G and B alike; its default
        # basis spans the same spaces, which the K agreement below checks
        want = ("K", "rhs") + (("S", "g") if want_schur else ())
        with oracle.tri_test_basis(1 if tri else 0):
            o, x = oracle.with_extended(*args, schur=want_schur)
            moved = [oracle.assemble(args[0], args[1], nd, *args[3:], want=want)
                     for nd in rounding_copies(g["nodes"], copies, 7 + bi)]
        if tri:
            o0 = oracle.assemble(*args, want=("K",))
            eb = rel_fro(o0["K"], o["K"])
            print(f"  oracle K, two triangle test bases: {eb:.2e}")
            assert eb < 1e-8
        pairs = [("K", "K"), ("rhs", "rhs")] + ([("schur", "S"), ("schur_rhs", "g")] if want_schur else [])
        for gk, ok in pairs:
            floor = max([rel_fro(o[ok], x[ok])] + [rel_fro(m[ok], x[ok]) for m in moved])
            err = rel_fro(res[gk], x[ok])
            print(f"  {gk:10s} gpu-vs-80bit {err:.2e} oracle-vs-80bit {floor:.2e} gpu-vs-oracle "
                  f"{rel_fro(res[gk], o[ok]):.2e}")
            assert err <= max(1e-12, 4.0 * floor), (g["topology"], g["k"], g["kt"], gk)
            worst = max(worst, err)
    A.close()
    print(f"worst {worst:.2e}")


def test_c3_every_element_type_matches_oracle():
    """All 18 (p, dp) quad types and the triangle types of a small graded mesh
    (75 cells, eps down to 1e-4) against the oracle, Schur complement included."""
    f, groups, load = c3_case(3, 3)
    assert sum(g["topology"] == capi.QUAD for g in groups) == 18
    check_types_against_oracle(f, groups, load)


def test_c3_finest_cells_of_every_type_match_oracle():
    """The full 49,424-cell graded mesh (10 levels): the two smallest cells of
    every element type (h down to 5.7e-5 with eps down to 1e-4, where the
    oracle's own FP64 distance from the 80-bit evaluation of K reaches 1e-4)."""
    f, groups, load = c3_case(16, 10, tri_cols=4)
    sel = []
    for g in groups:
        nd = g["nodes"]
        e1, e2 = nd[:, 1] - nd[:, 0], nd[:, -1] - nd[:, 0]
        idx = np.sort(np.argsort(np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]))[:2])
        sel.append(dict(g, nodes=nd[idx], parities=g["parities"][idx], coefs=g["coefs"][idx], ncells=len(idx)))
    check_types_against_oracle(f, sel, load)


@pytest.mark.parametrize("form", ["cd", "poisson"])
def test_triangle_element_types_match_oracle(form):
    """Triangles against the oracle. The oracle's default triangle test basis
    (unnormalised Koornwinder, interleaved RT, vertex complement) differs from
    the library's (orthonormal Dubiner by degree, centroid complement): its K
    agreeing to 1e-8 with the oracle's K in the library's basis family shows both
    span P_kt and RT_kt (K of other spaces differs at O(1)); the tight check is
    against the latter. The field basis is nodal on warp-and-blend nodes in both."""
    nodes, par = dpg.mesh_quad(3, 3, 0.1, 21)
    tn, tp = split_quads(nodes, par)
    C = tn.shape[0]
    f = dpg.form(form, 2)
    coefs = None
    if form == "cd":
        eps = 10.0 ** dpg.rng_uniform(5, -4.0, -2.0, C)
        coefs = np.stack([eps, np.full(C, 0.6), np.full(C, 0.8)], axis=1)
    load = dpg.load_sine_modes(2, 1, 8, 6, 0xC0FFEE + 101)
    groups = [dict(topology=capi.TRI, k=k, kt=kt, bucket=dpg.bucket(capi.TRI, k, kt), nodes=tn, parities=tp,
                   coefs=coefs, ncells=C) for k, kt in [(2, 2), (2, 4), (3, 5), (4, 4), (5, 8), (7, 9)]]
    check_types_against_oracle(f, groups, load)


def test_tile_edges_give_identical_results_and_autotune_picks_one():
    """Dense-G tile edges 64 and 32 give bit-identical results; the opt-in
    sum-factorised G agrees with the dense G to rounding (its K does not: see
    the accuracy note on dpg_tiles.sumfact)."""
    f, groups, load = c3_case(2, 3)
    sel = [groups[0], groups[-1]]  # largest and smallest element type
    A = dpg.LocalAssembler(f, [g["bucket"] for g in sel])
    for j, g in enumerate(sel):
        args = (dev(g["nodes"]), dev(g["parities"]), dev(g["coefs"]), load)
        assert A.tiles(j) == (64, 64, 64, 0)
        A.set_tiles(j, 64, 64, 64, 1)
        sfr, _ = A.local_stiffness_and_rhs(j, *args, want_schur=True, want_G=True)
        sfr = {key: v.cpu().numpy() for key, v in sfr.items()}
        A.set_tiles(j, 64, 64, 64, 0)
        a, _ = A.local_stiffness_and_rhs(j, *args, want_schur=True, want_G=True)
        a = {key: v.cpu().numpy() for key, v in a.items()}
        A.set_tiles(j, 32, 32, 32, 0)
        b, _ = A.local_stiffness_and_rhs(j, *args, want_schur=True, want_G=True)
        b = {key: v.cpu().numpy() for key, v in b.items()}
        A.set_tiles(j, 64, 64, 128, 0)  # 128 x 64 panel-update tile
        c, _ = A.local_stiffness_and_rhs(j, *args, want_schur=True, want_G=True)
        c = {key: v.cpu().numpy() for key, v in c.items()}
        for key in ("K", "rhs", "schur", "schur_rhs", "G"):
            assert np.array_equal(a[key], b[key]), key
        # the fused panel update and solve (update tile 128) solves blockwise:
        # the same G, the other outputs under the floor rule
        assert np.array_equal(a["G"], c["G"])
        o, x = oracle.with_extended(oracle.form_like(f), g["bucket"], g["nodes"], g["parities"], g["coefs"], load,
                                    schur=True)
        for key, ok in (("K", "K"), ("rhs", "rhs"), ("schur", "S"), ("schur_rhs", "g")):
            floor = rel_fro(o[ok], x[ok])
            assert rel_fro(c[key], x[ok]) <= max(1e-12, 4.0 * floor), key
        eg = rel_fro(sfr["G"], a["G"])
        ek = rel_fro(sfr["K"], a["K"])
        print(f"type {g['topology'], g['k'], g['kt']}: sum-factorised vs dense G {eg:.2e}, K {ek:.2e}")
        assert eg < 1e-14
        t = A.autotune(j, *args, sample=g["ncells"])
        print(f"  autotuned {t}")
        assert t[0] in (32, 64) and t[1] in (32, 64) and t[2] in (32, 64, 128) and t[3] in (0, 1)
        if t[3]:  # the sum-factorised G was taken: its outputs under the same bar
            e, _ = A.local_stiffness_and_rhs(j, *args, want_schur=True)
            e = {key: v.cpu().numpy() for key, v in e.items()}
            for key, ok in (("K", "K"), ("rhs", "rhs"), ("schur", "S"), ("schur_rhs", "g")):
                assert rel_fro(e[key], x[ok]) <= max(1e-12, 4.0 * rel_fro(o[ok], x[ok])), key
    with pytest.raises(capi.DPGError):
        A.set_tiles(0, 48, 64, 64)
    with pytest.raises(capi.DPGError):
        A.set_tiles(0, 128, 64, 64)  # the 128 tile is for the updates only
    A.close()
    At = dpg.LocalAssembler(f, [dpg.bucket(capi.TRI, 3, 4)])
    assert At.tiles(0) == (64, 64, 64, 0)
    with pytest.raises(capi.DPGError):
        At.set_tiles(0, 64, 64, 64, 1)
    At.close()


def test_graph_replays_match_per_bucket_launches():
    """One captured launch over all element types of a quad+triangle hp mesh
    (3 lanes, a workspace small enough that large types run in several
    batches) equals the per-bucket plain launches bit for bit, on every
    replay, with eager work between the replays."""
    f, groups, load = c3_case(4, 4, tri_cols=2)
    assert any(g["topology"] == capi.TRI for g in groups)
    buckets = [g["bucket"] for g in groups]
    A = dpg.LocalAssembler(f, buckets)
    ins = [(dev(g["nodes"]), dev(g["parities"]), dev(g["coefs"])) for g in groups]
    ref = []
    for bi, (n_, p_, c_) in enumerate(ins):
        out, nbad = A.local_stiffness_and_rhs(bi, n_, p_, c_, load, want_schur=True)
        assert nbad == 0
        ref.append({k: v.clone() for k, v in out.items()})
    torch.cuda.synchronize()
    # separate plan: the graph's workspace budget forces several batches
    Ag = dpg.LocalAssembler(f, buckets, workspace_bytes=48 * 2 ** 20)
    items = [dict(bucket=bi, nodes=n_, parities=p_, coefs=c_, load=load, want_schur=True)
             for bi, (n_, p_, c_) in enumerate(ins)]
    G = dpg.AssemblyGraph(Ag, items, lanes=3)
    print(f"graph workspace {G.workspace_bytes() / 2**20:.1f} MiB", flush=True)
    for rep in range(3):
        for o in G.outputs():
            for key in ("K", "rhs", "schur", "schur_rhs"):
                o[key].fill_(float("nan"))
            o["info"].fill_(-7)
        # unrelated eager work on the same device between replays
        junk, _ = A.local_stiffness_and_rhs(0, *ins[0], load)
        torch.cuda.synchronize()
        G.launch()
        torch.cuda.synchronize()
        for bi, o in enumerate(G.outputs()):
            assert torch.all(o["info"] == 0), (rep, bi)
            for key in ("K", "rhs", "schur", "schur_rhs"):
                assert torch.equal(o[key], ref[bi][key]), (rep, bi, key)
        assert torch.equal(junk["K"], ref[0]["K"])
    G.close()
    A.close()
    Ag.close()


def test_graph_writes_lapack_info_on_device():
    """INFO inside the graph: an IP with only the v term leaves G singular in the
    tau block, so every cell reports the 1-based column of its first tau dof."""
    f, groups, load = c3_case(2, 3)
    g = [x for x in groups if x["topology"] == capi.QUAD][-1]
    lt = f.ipterms[0]
    lt.n = 1
    lt.s[0].test_var, lt.s[0].test_op = 0, capi.OP_VALUE
    lt.s[0].coef.scale, lt.s[0].coef.coef_slot, lt.s[0].coef.coef_fn, lt.s[0].coef.normal_comp = 1.0, -1, 0, -1
    f.n_ipterms = 1
    A = dpg.LocalAssembler(f, [g["bucket"]])
    args = (dev(g["nodes"]), dev(g["parities"]), dev(g["coefs"]))
    G = dpg.AssemblyGraph(A, [dict(bucket=0, nodes=args[0], parities=args[1], coefs=args[2], load=load)], lanes=1)
    G.launch()
    torch.cuda.synchronize()
    info = G.outputs()[0]["info"].cpu().numpy()
    _, nbad = A.local_stiffness_and_rhs(0, *args, load)
    nv = (g["kt"] + 1) ** 2
    assert nbad == g["ncells"] and np.all(info == nv + 1)
    G.close()
    A.close()


@pytest.mark.parametrize("topo", [capi.QUAD, capi.TRI])
def test_c3_energy_error_vanishes_for_exact_linear_solution(topo):
    """GPU energy error of the interpolated linear solution (ConvectionDiffusion,
    eps down to 1e-4) on quads and triangles: the invariant the oracle tests pin."""
    nodes, par = dpg.mesh_quad(3, 3, 0.1, 17)
    if topo == capi.TRI:
        nodes, par = split_quads(nodes, par)
    C = nodes.shape[0]
    k, kt = 3, 4
    a = (0.4, -0.9, 1.3)
    beta = (np.cos(0.3), np.sin(0.3))
    eps = 10.0 ** dpg.rng_uniform(3, -4.0, -2.0, C)
    coefs = np.stack([eps, np.full(C, beta[0]), np.full(C, beta[1])], axis=1)
    ld = capi.LoadDesc()
    ld.ncomp = 1
    ld.constant[0] = beta[0] * a[1] + beta[1] * a[2]
    u = linear_trial_coefficients(nodes, par, k, a, eps=eps, beta=beta)
    f = dpg.form("cd", 2)
    A = dpg.LocalAssembler(f, [dpg.bucket(topo, k, kt)])
    eta2 = A.energy_error(0, dev(nodes), dev(par), dev(u), dev(coefs), ld).cpu().numpy()
    out, _ = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), dev(coefs), ld)
    K = out["K"].cpu().numpy()
    scale = np.einsum("cm,cmn,cn->c", u, K, u)
    print(f"max eta2/u^T K u {np.max(np.abs(eta2) / scale):.2e}")
    assert np.all(np.abs(eta2) <= 1e-12 * scale)
    A.close()
