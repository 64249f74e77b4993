"""CPU checks of the C-ABI library: it loads, exports every declared entry
point, and its host-side parts (formulations, graph norm, element sizes,
seeded inputs, argument checking) agree with the oracle and the reference."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

from camellia2_amd import capi, dpg
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    src = open(os.path.join(ROOT, "include", "dpg_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dpg_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_entry_points()
    assert len(names) >= 17
    for n in names:
        assert hasattr(L, n), n
    assert sorted(capi.EXPORTS) == names
    assert b"gfx950" in L.dpg_version()


def _ipterms(f):
    out = []
    for i in range(f.n_ipterms):
        lt = f.ipterms[i]
        out.append(sorted((lt.s[k].test_var, lt.s[k].test_op, lt.s[k].coef.scale, lt.s[k].coef.coef_slot,
                           lt.s[k].coef.coef_fn) for k in range(lt.n)))
    return out


def _bterms(f):
    return sorted((f.bterms[i].trial_var, f.bterms[i].trial_comp, f.bterms[i].test_var, f.bterms[i].test_op,
                   f.bterms[i].coef.scale, f.bterms[i].coef.coef_slot, f.bterms[i].coef.coef_fn,
                   f.bterms[i].coef.normal_comp) for i in range(f.n_bterms))


@pytest.mark.parametrize("name,dim", [("poisson", 2), ("poisson", 3), ("cdr", 2), ("cdr", 3), ("stokes", 2), ("cd", 2)])
def test_formulations_match_oracle_tables(name, dim):
    f = dpg.form(name, dim)
    fo = oracle.form(name, dim)
    assert _bterms(f) == _bterms(fo)
    assert _ipterms(f) == _ipterms(fo)
    assert list(f.trial_kind[:f.n_trial]) == list(fo.trial_kind[:fo.n_trial])
    assert list(f.test_space[:f.n_test]) == list(fo.test_space[:fo.n_test])
    g = dpg.graph_norm(dpg.form(name, dim))
    assert _ipterms(g) == _ipterms(fo)


@pytest.mark.parametrize("name,dim,topo,k,kt,N,M,Q,Qs,QG", [
    ("poisson", 2, capi.QUAD, 3, 4, 65, 55, 16, 4, 25),      # C1
    ("stokes", 2, capi.QUAD, 4, 5, 228, 184, 25, 5, 36),     # C2
    ("poisson", 3, capi.HEX, 4, 5, 666, 502, 125, 25, 216),  # C4
    ("cdr", 2, capi.QUAD, 3, 4, 65, 55, 16, 4, 25),          # C5
    ("cd", 2, capi.QUAD, 7, 9, 280, 207, 81, 9, 100),        # C3 largest quad type
    ("cd", 2, capi.TRI, 7, 9, 154, 129, 100, 9, 121),        # C3 largest triangle type
    ("cd", 2, capi.TRI, 2, 2, 14, 24, 16, 3, 16),            # C3 smallest triangle type
])
def test_bucket_sizes(name, dim, topo, k, kt, N, M, Q, Qs, QG):
    f = dpg.form(name, dim)
    b = dpg.bucket(topo, k, kt)
    s = dpg.sizes(f, b)
    assert (s.N, s.M, s.Q, s.Qs, s.QG) == (N, M, Q, Qs, QG)
    so = oracle.sizes(oracle.form(name, dim), b)
    assert (so["N"], so["M"], so["F"]) == (s.N, s.M, s.F)
    assert s.T == s.M - s.F


def test_mesh_is_seeded_and_positively_oriented():
    a, pa = dpg.mesh_quad(8, 8, 0.1, 5)
    b, pb = dpg.mesh_quad(8, 8, 0.1, 5)
    c, _ = dpg.mesh_quad(8, 8, 0.1, 6)
    assert np.array_equal(a, b) and np.array_equal(pa, pb) and not np.array_equal(a, c)
    # CCW: positive signed area of every quad
    x, y = a[..., 0], a[..., 1]
    area = 0.5 * np.sum(x * np.roll(y, -1, 1) - np.roll(x, -1, 1) * y, axis=1)
    assert area.min() > 0 and abs(area.sum() - 1.0) < 1e-12
    # interior sides: one cell +1, its neighbour -1
    nx = 8
    for cidx in range(64):
        i, j = cidx % nx, cidx // nx
        if i + 1 < nx:
            assert pa[cidx, 1] == 1 and pa[cidx + 1, 3] == -1
        if j + 1 < nx:
            assert pa[cidx, 2] == 1 and pa[cidx + nx, 0] == -1
    h, ph = dpg.mesh_hex(3, 3, 3, 0.05, 9)
    assert h.shape == (27, 8, 3) and ph.shape == (27, 6)
    assert np.all(ph[0] == 1)


def test_rng_statistics():
    u = dpg.rng_uniform(1, -1.0, 1.0, 200000)
    n = dpg.rng_normal(2, 200000)
    assert abs(u.mean()) < 0.01 and abs(u.var() - 1 / 3) < 0.01 and u.min() >= -1 and u.max() < 1
    assert abs(n.mean()) < 0.01 and abs(n.var() - 1.0) < 0.02


def test_bad_arguments_return_errors_not_crashes():
    L = capi.lib()
    f = capi.FormDesc()
    assert L.dpg_form_poisson_uw(4, ct.byref(f)) < 0
    assert b"spatial_dim" in L.dpg_last_error()
    assert L.dpg_mesh_quad(0, 4, 0.1, 1, None, None) < 0
    f = dpg.form("poisson", 2)
    info = capi.BucketInfo()
    assert L.dpg_bucket_sizes(ct.byref(f), ct.byref(capi.BucketDesc(capi.HEX, 3, 4, 0)), ct.byref(info)) < 0
    f.bterms[0].test_op = capi.OP_DOT_NORMAL  # (u, tau.n) puts a field on the skeleton
    f.bterms[0].trial_var = 0
    assert L.dpg_bucket_sizes(ct.byref(f), ct.byref(capi.BucketDesc(capi.QUAD, 3, 4, 0)), ct.byref(info)) < 0


def test_plan_without_gpu_reports_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = capi.lib()
    f = dpg.form("poisson", 2)
    arr = (capi.BucketDesc * 1)(capi.BucketDesc(capi.QUAD, 3, 4, 0))
    plan = ct.c_void_p()
    rc = L.dpg_plan_create(ct.byref(f), arr, 1, None, ct.byref(plan))
    assert rc < 0 and not plan.value
    assert len(L.dpg_last_error()) > 0


def test_graded_mesh_is_a_consistent_hp_mesh():
    n0, levels = 4, 5
    nodes, par, lvl = dpg.mesh_quad_graded(n0, levels, 0.05, 11)
    C = nodes.shape[0]
    assert C == n0 * n0 + 3 * n0 * (2 ** levels - 1)
    assert np.array_equal(np.bincount(lvl)[1:], [n0 * 2 ** l for l in range(1, levels)] + [n0 * 2 ** (levels + 1)])
    n2, p2, _ = dpg.mesh_quad_graded(n0, levels, 0.05, 11)
    assert np.array_equal(nodes, n2) and np.array_equal(par, p2)
    x, y = nodes[..., 0], nodes[..., 1]
    area = 0.5 * np.sum(x * np.roll(y, -1, 1) - np.roll(x, -1, 1) * y, axis=1)
    assert area.min() > 0 and abs(area.sum() - 1.0) < 1e-13  # a tiling: no gaps, no overlaps
    assert np.all(np.diff(lvl) >= 0)  # coarse cells first

    def key(pt):
        return (round(pt[0], 12), round(pt[1], 12))
    edges = {}
    for c in range(C):
        for s in range(4):
            a, b = nodes[c, s], nodes[c, (s + 1) % 4]
            edges.setdefault(frozenset([key(a), key(b)]), []).append((c, s, a, b))
    on_boundary = lambda a, b: (abs(a[0] - b[0]) < 1e-14 and abs(a[0]) % 1.0 < 1e-14) or \
        (abs(a[1] - b[1]) < 1e-14 and abs(a[1]) % 1.0 < 1e-14)
    coarse_mid, fine = {}, []
    for e, cs in edges.items():
        if len(cs) == 2:  # conforming interior side: opposite parities
            assert par[cs[0][0], cs[0][1]] == -par[cs[1][0], cs[1][1]]
        else:
            c, s, a, b = cs[0]
            if on_boundary(a, b):
                assert par[c, s] == 1
            else:
                fine.append(cs[0])
                coarse_mid[key(0.5 * (a + b))] = cs[0]
    nhang = 0
    for c, s, a, b in fine:
        for pt in (a, b):
            if key(pt) in coarse_mid and coarse_mid[key(pt)][0] != c:
                cc, cs_, _, _ = coarse_mid[key(pt)]
                assert lvl[cc] == lvl[c] - 1 and par[cc, cs_] == 1 and par[c, s] == -1
                nhang += 1
    assert nhang == 2 * n0 * (2 ** levels - 1)  # two fine sides per refined coarse side


def test_form_too_large_for_the_operand_lds_is_refused():
    """The operand kernels stage the weights of 8 points x groups x test
    variables x prims in LDS: a form beyond 64 KiB is refused when it is
    validated, with a message, not at launch."""
    f = dpg.form("poisson", 3)
    f.n_test = 16
    for v in range(2, 16):
        f.test_space[v] = capi.HGRAD
    f.n_trial = 16
    for v in range(4, 16):
        f.trial_kind[v] = capi.FIELD_VECTOR
    with pytest.raises(capi.DPGError, match="LDS"):
        dpg.sizes(f, dpg.bucket(capi.HEX, 2, 3))
