"""hp meshes on the GPU: every (p, dp) element type of the graded
convection-diffusion mesh against the oracle, GEMM tile choices, and the
all-bucket hipGraph launch against per-bucket plain launches (bit for bit)."""
import numpy as np
import pytest

from camellia2_amd import capi, dpg
from oracle import oracle

from helpers import cd_hp, hp_groups, rel_fro

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def c3_case(n0, levels, **kw):
    f, nodes, par, coefs, load, p, dp, lvl = cd_hp(n0, levels, **kw)
    groups = hp_groups(p, dp)
    buckets = [dpg.bucket(capi.QUAD, k, kt) for (k, kt) in groups]
    return f, nodes, par, coefs, load, groups, buckets


def test_c3_every_element_type_matches_oracle():
    """All 18 (p, dp) element types of a small graded mesh (46 cells, eps down
    to 1e-4) against the oracle: K, rhs and the Schur complement within
    max(1e-12, 4 x the oracle's own distance from an 80-bit evaluation)."""
    f, nodes, par, coefs, load, groups, buckets = c3_case(2, 3)
    assert len(buckets) == 18
    A = dpg.LocalAssembler(f, buckets)
    worst = 0.0
    for bi, ((k, kt), idx) in enumerate(groups.items()):
        if len(idx) == 0:
            continue
        print(f"bucket k={k} kt={kt} cells={len(idx)}", flush=True)
        out, nbad = A.local_stiffness_and_rhs(bi, dev(nodes[idx]), dev(par[idx]), dev(coefs[idx]), load,
                                              want_schur=True)
        torch.cuda.synchronize()
        g = {key: v.cpu().numpy() for key, v in out.items()}
        assert nbad == 0 and np.all(g["info"] == 0)
        o, x = oracle.with_extended(oracle.form_like(f), buckets[bi], nodes[idx], par[idx], coefs[idx], load,
                                    schur=True)
        for gk, ok in (("K", "K"), ("rhs", "rhs"), ("schur", "S"), ("schur_rhs", "g")):
            floor = rel_fro(o[ok], x[ok])
            err = rel_fro(g[gk], x[ok])
            print(f"  {gk:10s} gpu-vs-80bit {err:.2e} oracle-vs-80bit {floor:.2e}")
            assert err <= max(1e-12, 4.0 * floor), (k, kt, gk)
            worst = max(worst, err)
    A.close()
    print(f"worst {worst:.2e}")


def test_tile_edges_give_identical_results_and_autotune_picks_one():
    f, nodes, par, coefs, load, groups, buckets = c3_case(2, 3)
    keys = list(groups)
    sel = [0, len(keys) - 1]  # largest and smallest element type
    A = dpg.LocalAssembler(f, [buckets[i] for i in sel])
    for j, i in enumerate(sel):
        idx = groups[keys[i]]
        args = (dev(nodes[idx]), dev(par[idx]), dev(coefs[idx]), load)
        assert A.tiles(j) == (64, 64, 64)
        a, _ = A.local_stiffness_and_rhs(j, *args, want_schur=True)
        a = {key: v.cpu().numpy() for key, v in a.items()}
        A.set_tiles(j, 32, 32, 32)
        b, _ = A.local_stiffness_and_rhs(j, *args, want_schur=True)
        b = {key: v.cpu().numpy() for key, v in b.items()}
        for key in ("K", "rhs", "schur", "schur_rhs"):
            assert np.array_equal(a[key], b[key]), key
        t = A.autotune(j, *args, sample=len(idx))
        print(f"bucket {keys[i]}: autotuned tiles {t}")
        assert all(v in (32, 64) for v in t)
    with pytest.raises(capi.DPGError):
        A.set_tiles(0, 48, 64, 64)
    A.close()


def test_graph_replays_match_per_bucket_launches():
    """One captured launch over all element types (3 lanes, a workspace small
    enough that large types run in several batches) equals the per-bucket
    plain launches bit for bit, on every replay, with eager work between."""
    f, nodes, par, coefs, load, groups, buckets = c3_case(4, 4)
    A = dpg.LocalAssembler(f, buckets)
    ref = []
    for bi, idx in enumerate(groups.values()):
        out, nbad = A.local_stiffness_and_rhs(bi, dev(nodes[idx]), dev(par[idx]), dev(coefs[idx]), load,
                                              want_schur=True)
        assert nbad == 0
        ref.append({k: v.clone() for k, v in out.items()})
    torch.cuda.synchronize()
    # separate plan: the graph's workspace budget forces several batches
    Ag = dpg.LocalAssembler(f, buckets, workspace_bytes=48 * 2 ** 20)
    items = [dict(bucket=bi, nodes=dev(nodes[idx]), parities=dev(par[idx]), coefs=dev(coefs[idx]), load=load,
                  want_schur=True) for bi, idx in enumerate(groups.values())]
    G = dpg.AssemblyGraph(Ag, items, lanes=3)
    print(f"graph workspace {G.workspace_bytes() / 2**20:.1f} MiB", flush=True)
    for rep in range(3):
        for o in G.outputs():
            for key in ("K", "rhs", "schur", "schur_rhs"):
                o[key].fill_(float("nan"))
            o["info"].fill_(-7)
        # unrelated eager work on the same device between replays
        junk, _ = A.local_stiffness_and_rhs(0, items[0]["nodes"], items[0]["parities"], items[0]["coefs"], load)
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
    f, nodes, par, coefs, load, groups, buckets = c3_case(2, 3)
    (k, kt), idx = list(groups.items())[-1]
    lt = f.ipterms[0]
    lt.n = 1
    lt.s[0].test_var, lt.s[0].test_op = 0, capi.OP_VALUE
    lt.s[0].coef.scale, lt.s[0].coef.coef_slot, lt.s[0].coef.coef_fn, lt.s[0].coef.normal_comp = 1.0, -1, 0, -1
    f.n_ipterms = 1
    A = dpg.LocalAssembler(f, [buckets[-1]])
    G = dpg.AssemblyGraph(A, [dict(bucket=0, nodes=dev(nodes[idx]), parities=dev(par[idx]), coefs=dev(coefs[idx]),
                                   load=load)], lanes=1)
    G.launch()
    torch.cuda.synchronize()
    info = G.outputs()[0]["info"].cpu().numpy()
    _, nbad = A.local_stiffness_and_rhs(0, dev(nodes[idx]), dev(par[idx]), dev(coefs[idx]), load)
    nv = (kt + 1) ** 2
    assert nbad == len(idx) and np.all(info == nv + 1)
    G.close()
    A.close()


def test_c3_energy_error_vanishes_for_exact_linear_solution():
    """GPU energy error of the interpolated linear solution on the hp mesh's
    lowest-order element type (the invariant the oracle test pins)."""
    from test_oracle import _nodal_trial_coefficients_linear
    f, nodes, par, coefs, load, groups, buckets = c3_case(2, 3)
    k, kt = 3, 4
    idx = np.arange(nodes.shape[0])
    a = (0.4, -0.9, 1.3)
    beta = (coefs[0, 1], coefs[0, 2])
    ld = capi.LoadDesc()
    ld.ncomp = 1
    ld.constant[0] = beta[0] * a[1] + beta[1] * a[2]
    u = _nodal_trial_coefficients_linear(nodes, par, k, a, eps=coefs[:, 0], beta=beta)
    A = dpg.LocalAssembler(f, [dpg.bucket(capi.QUAD, k, kt)])
    eta2 = A.energy_error(0, dev(nodes[idx]), dev(par[idx]), dev(u), dev(coefs[idx]), ld).cpu().numpy()
    out, _ = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), dev(coefs), ld)
    K = out["K"].cpu().numpy()
    scale = np.einsum("cm,cmn,cn->c", u, K, u)
    print(f"max eta2/u^T K u {np.max(np.abs(eta2) / scale):.2e}")
    assert np.all(np.abs(eta2) <= 1e-12 * scale)
    A.close()
