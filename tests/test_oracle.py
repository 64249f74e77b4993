"""Pins the oracle (CPU restatement of the reference's local assembly) against
the reference's own golden vectors and against what the mathematics fixes."""
import json
import os

import numpy as np
import pytest

from camellia2_amd import capi, dpg
from oracle import oracle

from helpers import (HEX_FACES, HEX_VERTS, linear_trial_coefficients, nodal_trial_coefficients, side_points,
                     split_quads, tensor_nodes)


def load(golden_dir, name):
    with open(os.path.join(golden_dir, name)) as f:
        return json.load(f)


def coef(scale=1.0, slot=-1, fn=0, normal=-1):
    c = capi.Coef()
    c.scale, c.coef_slot, c.coef_fn, c.normal_comp = scale, slot, fn, normal
    return c


def add_ip(f, summands):
    lt = f.ipterms[f.n_ipterms]
    lt.n = len(summands)
    for i, (var, op, c) in enumerate(summands):
        lt.s[i].test_var, lt.s[i].test_op, lt.s[i].coef = var, op, c
    f.n_ipterms += 1


def analytic_boundary_form():
    """b(uhat, v) = <uhat n, grad v>, IP = (v,v) + (grad v, grad v)
    (TestBilinearFormAnalyticBoundaryIntegral.h:42-50, MathInnerProduct.cpp:17-60)."""
    f = capi.FormDesc()
    f.spatial_dim, f.n_trial, f.n_test = 2, 1, 1
    f.trial_kind[0] = capi.TRACE
    f.test_space[0] = capi.HGRAD
    for d in range(2):
        b = f.bterms[f.n_bterms]
        b.trial_var, b.trial_comp, b.test_var, b.test_op = 0, 0, 0, capi.OP_DX + d
        b.coef = coef(normal=d)
        f.n_bterms += 1
    for op in (capi.OP_VALUE, capi.OP_DX, capi.OP_DY):
        add_ip(f, [(0, op, coef())])
    return f


REF_QUAD = np.array([[[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]]])


@pytest.mark.parametrize("case", range(4))
def test_factored_cholesky_kats(golden_dir, case):
    """BFTests.cpp:28-72,136-381 (TEST_FLOATING_EQUALITY, tol 1e-14)."""
    kats = load(golden_dir, "factored_cholesky_kats.json")
    c = kats["cases"][case]
    K, rhs, info = oracle.factored_cholesky_solve(c["G_colmajor"], c["B_colmajor"], c["l"], c["N"], c["M"])
    assert info == 0
    Ke, re = np.array(c["K"]), np.array(c["rhs"])
    tol = kats["tol"]
    mask = np.abs(Ke) > tol
    assert np.all(np.abs(K[mask] - Ke[mask]) <= tol * np.abs(Ke[mask]) * 4)
    assert np.all(np.abs(K[~mask]) < tol)
    maskr = np.abs(re) > tol
    assert np.all(np.abs(rhs[maskr] - re[maskr]) <= tol * np.abs(re[maskr]) * 4)
    assert np.all(np.abs(rhs[~maskr]) < tol)


def test_analytic_boundary_integral_cubic(golden_dir):
    """DPGTests.cpp:966-1098 with the Mathematica values, tol 3.76e-14."""
    fx = load(golden_dir, "analytic_boundary_integral_cubic.json")
    f = analytic_boundary_form()
    bk = capi.BucketDesc(capi.QUAD, 3, 3, 0)
    o = oracle.assemble(f, bk, REF_QUAD, np.ones((1, 4), np.int8), want=("B", "G", "K"))
    tol = fx["tol"]
    B, G, K = o["B"][0], o["G"][0], o["K"][0]
    assert np.abs(B.T - np.array(fx["pre_stiffness"]["nonconforming"])).max() < tol
    assert np.abs(G - np.array(fx["ip_matrix"]["only"])).max() < tol * 1.01
    W = np.linalg.solve(G, B.T).T
    assert np.abs(W - np.array(fx["opt_test_weights"]["nonconforming"])).max() < tol
    assert np.abs(K - np.array(fx["final_stiffness"]["nonconforming"])).max() < tol


def test_analytic_boundary_integral_conforming(golden_dir):
    """Conforming traces identify the vertex dofs of neighbouring sides
    (TestBilinearFormAnalyticBoundaryIntegral.h:60-81): K_conf = P^T K P."""
    fx = load(golden_dir, "analytic_boundary_integral_cubic.json")
    f = analytic_boundary_form()
    bk = capi.BucketDesc(capi.QUAD, 3, 3, 0)
    o = oracle.assemble(f, bk, REF_QUAD, np.ones((1, 4), np.int8), want=("B", "K"))
    # side s: local dofs 4s..4s+3; last dof of side s == first dof of side s+1
    P = np.zeros((16, 12))
    for s in range(4):
        for a in range(4):
            g = (3 * s + a) % 12
            P[4 * s + a, g] = 1.0
    Bc = P.T @ o["B"][0]
    Kc = P.T @ o["K"][0] @ P
    assert np.abs(Bc.T - np.array(fx["pre_stiffness"]["conforming"])).max() < fx["tol"]
    assert np.abs(Kc - np.array(fx["final_stiffness"]["conforming"])).max() < fx["tol"]


def test_math_inner_product_c1(golden_dir):
    """DPGTests.cpp:755-813: C1 quad (vertex ordering) IP matrix."""
    fx = load(golden_dir, "math_inner_product_c1.json")
    f = analytic_boundary_form()
    bk = capi.BucketDesc(capi.QUAD, 1, 1, 0)
    o = oracle.assemble(f, bk, REF_QUAD, np.ones((1, 4), np.int8), want=("G",))
    perm = [0, 1, 3, 2]  # vertex order -> tensor (x fastest) order
    G = o["G"][0][np.ix_(perm, perm)]
    assert np.abs(G - np.array(fx["ip_matrix"])).max() < fx["tol"]


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11])
def test_gauss_rule_matches_numpy(n):
    x, w = oracle.gauss_rule(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.abs(x - xr).max() < 1e-14 and np.abs(w - wr).max() < 1e-14
    # exactness up to degree 2n-1
    for d in range(2 * n):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(np.dot(w, x ** d) - exact) < 1e-13


def test_gll_nodes_closed_form():
    assert np.allclose(oracle.gll_nodes(2), [-1, 0, 1], atol=1e-15)
    assert np.allclose(oracle.gll_nodes(3), [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1], atol=1e-15)
    assert np.allclose(oracle.gll_nodes(4), [-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1], atol=1e-15)
    for n in range(2, 10):
        x = oracle.gll_nodes(n)
        dP = np.polynomial.legendre.Legendre.basis(n).deriv()
        assert np.abs(dP(x[1:-1])).max() < 1e-10 * max(1.0, np.abs(dP.coef).max())


def _quad_mesh(nx, ny, jitter, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / nx
    X = np.zeros((ny + 1, nx + 1, 2))
    for j in range(ny + 1):
        for i in range(nx + 1):
            X[j, i] = (i * h, j / ny)
            if 0 < i < nx and 0 < j < ny:
                X[j, i] += rng.uniform(-jitter, jitter, 2) * h
    nodes = np.zeros((nx * ny, 4, 2))
    for j in range(ny):
        for i in range(nx):
            nodes[j * nx + i] = [X[j, i], X[j, i + 1], X[j + 1, i + 1], X[j + 1, i]]
    par = np.ones((nx * ny, 4), np.int8)
    return nodes, par


def test_graph_norm_poisson_terms():
    """BF.cpp:1039-1141 on PoissonFormulation: (div tau)^2 + |tau - grad v|^2 + v^2 + |tau|^2."""
    f = oracle.form("poisson", 2)
    got = []
    for i in range(f.n_ipterms):
        lt = f.ipterms[i]
        got.append(sorted((lt.s[k].test_var, lt.s[k].test_op, lt.s[k].coef.scale) for k in range(lt.n)))
    V, TAU = 0, 1
    want = [[(TAU, capi.OP_DIV, 1.0)],
            sorted([(TAU, capi.OP_X, 1.0), (V, capi.OP_DX, -1.0)]),
            sorted([(TAU, capi.OP_Y, 1.0), (V, capi.OP_DY, -1.0)]),
            [(V, capi.OP_VALUE, 1.0)], [(TAU, capi.OP_X, 1.0)], [(TAU, capi.OP_Y, 1.0)]]
    assert got == want


_nodal_trial_coefficients_linear = linear_trial_coefficients


@pytest.mark.parametrize("k,kt", [(2, 3), (3, 4)])
def test_poisson_linear_solution_has_zero_residual(k, kt):
    """A linear u lies in the trial space of a bilinear mesh; with f = lap u = 0
    its DPG residual l - B u and energy error vanish (the TracesUltraweak
    invariant of ConvectionDiffusionReactionFormulationTests.cpp:116-235)."""
    nodes, par = _quad_mesh(3, 3, 0.1, 7)
    par[:, 0] = -1  # parities flip the flux sign; the residual must still vanish
    f = oracle.form("poisson", 2)
    bk = capi.BucketDesc(capi.QUAD, k, kt, 0)
    a = (0.3, -1.2, 0.7)
    u = _nodal_trial_coefficients_linear(nodes, par, k, a)
    o = oracle.assemble(f, bk, nodes, par, want=("B", "l"))
    r = o["l"] - np.einsum("cmn,cm->cn", o["B"], u)
    scale = np.abs(o["B"]).max() * np.abs(u).max()
    assert np.abs(r).max() < 1e-13 * scale
    eta2 = oracle.energy_error(f, bk, nodes, par, u)
    assert np.abs(eta2).max() < 1e-24 * scale ** 2 * 1e4


@pytest.mark.parametrize("k,kt", [(2, 3), (3, 5), (2, 2)])
def test_convection_diffusion_linear_solution_has_zero_residual(k, kt):
    """ConvectionDiffusionFormulation with per-cell eps (entering as 1/eps in
    (sigma/eps, tau)) and constant beta: u linear, sigma = eps grad u and
    f = div(beta u - sigma) = beta . grad u (a constant load) give l = B u.
    kt = k is the dp = 1 case (fewer test than trial dofs)."""
    nodes, par = _quad_mesh(3, 3, 0.1, 5)
    par[:, 3] = -1
    C = nodes.shape[0]
    eps = 10.0 ** np.linspace(-4, -2, C)
    beta = (np.cos(0.7), np.sin(0.7))
    coefs = np.stack([eps, np.full(C, beta[0]), np.full(C, beta[1])], axis=1)
    a = (0.4, -0.9, 1.3)
    f = oracle.form("cd", 2)
    bk = capi.BucketDesc(capi.QUAD, k, kt, 0)
    load = capi.LoadDesc()
    load.ncomp = 1
    load.constant[0] = beta[0] * a[1] + beta[1] * a[2]
    u = _nodal_trial_coefficients_linear(nodes, par, k, a, eps=eps, beta=beta)
    o = oracle.assemble(f, bk, nodes, par, coefs=coefs, load=load, want=("B", "l"))
    r = o["l"] - np.einsum("cmn,cm->cn", o["B"], u)
    for c in range(C):
        scale = np.abs(o["B"][c]).max() * np.abs(u[c]).max()
        assert np.abs(r[c]).max() < 1e-12 * scale, c
    # the same u with eps entering as eps (not 1/eps) leaves a residual
    wrong = _nodal_trial_coefficients_linear(nodes, par, k, a, eps=1.0 / eps, beta=beta)
    rw = o["l"] - np.einsum("cmn,cm->cn", o["B"], wrong)
    assert np.abs(rw).max() > 1e-3


def test_poisson_residual_detects_wrong_parity():
    """The same check fails when the flux coefficients ignore the parity."""
    nodes, par = _quad_mesh(2, 2, 0.1, 3)
    par[:, 1] = -1
    f = oracle.form("poisson", 2)
    bk = capi.BucketDesc(capi.QUAD, 2, 3, 0)
    u = _nodal_trial_coefficients_linear(nodes, np.ones_like(par), 2, (0.0, 1.0, 0.5))
    o = oracle.assemble(f, bk, nodes, par, want=("B", "l"))
    r = o["l"] - np.einsum("cmn,cm->cn", o["B"], u)
    assert np.abs(r).max() > 1e-3


def test_K_symmetric_spd_and_schur_closed_form():
    nodes, par = _quad_mesh(2, 2, 0.1, 11)
    f = oracle.form("poisson", 2)
    bk = capi.BucketDesc(capi.QUAD, 3, 4, 0)
    o = oracle.assemble(f, bk, nodes, par, want=("B", "G", "l", "K", "rhs", "S", "g"))
    s = oracle.sizes(f, bk)
    F = s["F"]
    for c in range(nodes.shape[0]):
        K = o["K"][c]
        assert np.array_equal(K, K.T)
        G, B, l = o["G"][c], o["B"][c], o["l"][c]
        Kr = B @ np.linalg.solve(G, B.T)
        assert np.abs(K - Kr).max() < 1e-10 * np.abs(K).max()
        assert np.abs(o["rhs"][c] - B @ np.linalg.solve(G, l)).max() < 1e-10 * np.abs(K).max()
        KFF, KFT, KTT = K[:F, :F], K[:F, F:], K[F:, F:]
        S = KTT - KFT.T @ np.linalg.solve(KFF, KFT)
        g = o["rhs"][c][F:] - KFT.T @ np.linalg.solve(KFF, o["rhs"][c][:F])
        assert np.abs(o["S"][c] - S).max() < 1e-9 * np.abs(S).max()
        assert np.abs(o["g"][c] - g).max() < 1e-9 * max(1.0, np.abs(g).max())
        assert np.linalg.eigvalsh(G).min() > 0


def test_indefinite_gram_reports_info():
    """POTRF INFO semantics (BF.cpp:963-969): a non-SPD G reports the failing column."""
    G = np.eye(3)
    G[1, 1] = -1.0
    K, rhs, info = oracle.factored_cholesky_solve(G.ravel(order="F"), np.eye(3).ravel(), np.ones(3), 3, 3)
    assert info == 2


# ------------------------------------------------------------------ triangles
@pytest.mark.parametrize("n", [2, 3, 5, 8, 12])
def test_triangle_rule_is_exact(n):
    """n x n collapsed Gauss points integrate x^a y^b exactly for a+b <= 2n-2:
    the integral over the unit triangle is a! b! / (a+b+2)!."""
    from math import factorial
    xy, w = oracle.tri_rule(n)
    assert np.all(w > 0) and np.all(xy[:, 0] + xy[:, 1] < 1) and np.all(xy > 0)
    for a in range(2 * n - 1):
        for b in range(2 * n - 1 - a):
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            got = np.sum(w * xy[:, 0] ** a * xy[:, 1] ** b)
            assert abs(got - exact) <= 1e-14 * exact + 1e-17, (a, b)


@pytest.mark.parametrize("n", range(1, 10))
def test_triangle_warp_blend_nodes(n):
    """Lattice for n <= 2; edge nodes at the 1D GLL points on all three edges;
    the node set is invariant under the symmetries of the triangle; the
    interior node of n = 3 is the centroid."""
    xy = oracle.tri_nodes(n)
    assert xy.shape == ((n + 1) * (n + 2) // 2, 2)
    if n <= 2:
        lat = np.array([(i / n, j / n) for j in range(n + 1) for i in range(n + 1 - j)])
        assert np.abs(xy - lat).max() < 1e-15
    g = 0.5 * (oracle.gll_nodes(n) + 1.0)
    bary = np.stack([1 - xy[:, 0] - xy[:, 1], xy[:, 0], xy[:, 1]], axis=1)
    for k in range(3):  # edge opposite vertex k: bary[k] == 0
        on = np.abs(bary[:, k]) < 1e-14
        other = bary[on][:, (k + 1) % 3]
        assert np.abs(np.sort(other) - g).max() < 1e-14
    key = lambda b: sorted(tuple(np.round(r, 12)) for r in b)
    for perm in ([1, 0, 2], [0, 2, 1], [2, 1, 0], [1, 2, 0]):
        assert key(bary[:, perm]) == key(bary)
    if n == 3:
        assert np.abs(xy[5] - 1.0 / 3.0).max() < 1e-15
    # lattice order: row j (n+1-j nodes) from the left edge to the hypotenuse, x increasing
    i0 = 0
    for j in range(n + 1):
        row = xy[i0:i0 + n + 1 - j]
        assert abs(row[0, 0]) < 1e-14 and abs(row[-1].sum() - 1.0) < 1e-14
        assert np.all(np.diff(row[:, 0]) > 0)
        i0 += n + 1 - j


def test_triangle_gram_of_l2_product_is_diagonal():
    """IP = (v, v) on P_n of an affine triangle: the oracle's orthogonal test
    basis gives a diagonal G with G_ii = 2 |T| / (2 (2i+1)(i+j+1))."""
    f = capi.FormDesc()
    f.spatial_dim, f.n_trial, f.n_test = 2, 1, 1
    f.trial_kind[0] = capi.FIELD
    f.test_space[0] = capi.HGRAD
    add_ip(f, [(0, capi.OP_VALUE, coef())])
    nodes = np.array([[[0.1, 0.2], [0.9, 0.35], [0.3, 1.1]]])
    e1, e2 = nodes[0, 1] - nodes[0, 0], nodes[0, 2] - nodes[0, 0]
    area = 0.5 * abs(e1[0] * e2[1] - e1[1] * e2[0])
    n = 6
    G = oracle.assemble(f, capi.BucketDesc(capi.TRI, 2, n, 0), nodes, np.ones((1, 3), np.int8), want=("G",))["G"][0]
    want = [2 * area / (2 * (2 * i + 1) * (i + j + 1)) for i in range(n + 1) for j in range(n + 1 - i)]
    assert np.abs(G - np.diag(want)).max() < 1e-14 * max(want)


def _tri_mesh(nx, ny, jitter, seed):
    nodes, par = _quad_mesh(nx, ny, jitter, seed)
    return split_quads(nodes, par)


@pytest.mark.parametrize("k,kt", [(2, 3), (3, 5), (4, 4), (7, 9)])
def test_triangle_poisson_linear_solution_has_zero_residual(k, kt):
    """The linear-solution invariant on triangles: normals, side measures and
    parities of the three sides, the RT test space and the nodal field basis."""
    nodes, par = _tri_mesh(2, 2, 0.1, 3)
    f = oracle.form("poisson", 2)
    bk = capi.BucketDesc(capi.TRI, k, kt, 0)
    u = linear_trial_coefficients(nodes, par, k, (0.3, -1.2, 0.7))
    o = oracle.assemble(f, bk, nodes, par, want=("B", "l"))
    r = o["l"] - np.einsum("cmn,cm->cn", o["B"], u)
    scale = np.abs(o["B"]).max() * np.abs(u).max()
    assert np.abs(r).max() < 1e-12 * scale
    wrong = linear_trial_coefficients(nodes, np.ones_like(par), k, (0.3, -1.2, 0.7))
    assert np.abs(o["l"] - np.einsum("cmn,cm->cn", o["B"], wrong)).max() > 1e-3


@pytest.mark.parametrize("k,kt", [(2, 3), (5, 7), (3, 3)])
def test_triangle_convection_diffusion_linear_solution_has_zero_residual(k, kt):
    nodes, par = _tri_mesh(2, 2, 0.1, 9)
    C = nodes.shape[0]
    eps = 10.0 ** np.linspace(-4, -2, C)
    beta = (np.cos(2.1), np.sin(2.1))
    coefs = np.stack([eps, np.full(C, beta[0]), np.full(C, beta[1])], axis=1)
    a = (0.4, -0.9, 1.3)
    load = capi.LoadDesc()
    load.ncomp = 1
    load.constant[0] = beta[0] * a[1] + beta[1] * a[2]
    f = oracle.form("cd", 2)
    u = linear_trial_coefficients(nodes, par, k, a, eps=eps, beta=beta)
    o = oracle.assemble(f, capi.BucketDesc(capi.TRI, k, kt, 0), nodes, par, coefs=coefs, load=load,
                        want=("B", "l"))
    r = o["l"] - np.einsum("cmn,cm->cn", o["B"], u)
    for c in range(C):
        assert np.abs(r[c]).max() < 1e-12 * np.abs(o["B"][c]).max() * np.abs(u[c]).max(), c


def test_extended_accumulation_changes_only_rounding():
    """The 80-bit evaluation sums the same FP64 point values: G, B, l agree with
    the FP64 sums to rounding, and its K is within the FP64 floor of the oracle's."""
    nodes, par = _quad_mesh(2, 2, 0.1, 13)
    f = oracle.form("poisson", 2)
    bk = capi.BucketDesc(capi.QUAD, 3, 4, 0)
    load = capi.LoadDesc()
    load.ncomp, load.nmodes[0], load.amp[0][0] = 1, 1, 1.3
    load.k[0][0][0], load.k[0][0][1] = 2, 3
    o, x = oracle.with_extended(f, bk, nodes, par, load=load, schur=True)
    lib = oracle.lib()
    lib.oracle_set_extended_accumulation(1)
    try:
        ox = oracle.assemble(f, bk, nodes, par, load=load, want=("B", "G", "l"))
    finally:
        lib.oracle_set_extended_accumulation(0)
    for k in ("B", "G", "l"):
        d = np.abs(ox[k] - o[k]).max() / np.abs(o[k]).max()
        assert 0 < d < 1e-14, k  # different rounding, same sums
    for k in ("K", "rhs", "S", "g"):
        rel = np.linalg.norm(o[k] - x[k]) / np.linalg.norm(x[k])
        assert rel < 1e-11, k


# ---------------------------------------------------------------------------
# Manufactured solutions that pin the oracle's Stokes VGP, CDR-with-reaction and
# hex term tables: an exact solution that the trial space represents exactly
# has l - B u = 0 on every cell (the TracesUltraweak invariant of
# ConvectionDiffusionReactionFormulationTests.cpp:116-235; the steady Stokes
# consistency check of StokesVGPFormulationTests.cpp:141-146), and a change of
# one term's sign, of a coefficient function or of a side parity breaks it.

def _residual(f, bk, nodes, par, u, coefs=None, load=None):
    """per cell max |l - B u| / (max|B| max|u|)"""
    o = oracle.assemble(f, bk, nodes, par, coefs=coefs, load=load, want=("B", "l"))
    r = o["l"] - np.einsum("cmn,cm->cn", o["B"], u)
    scale = np.abs(o["B"]).max(axis=(1, 2)) * np.abs(u).max(axis=1)
    return np.abs(r).max(axis=1) / scale


def _const_load(values):
    load = capi.LoadDesc()
    load.ncomp = len(values)
    for i, v in enumerate(values):
        load.constant[i] = v
    return load


def _stokes_solution(mu):
    """divergence-free linear velocity, linear pressure (StokesVGPFormulation.cpp:156-408):
    sigma = mu grad u, t_in = p n_i - sigma_i . n, f = grad p"""
    a1, a2, b11, b12, b21 = 0.4, -0.7, 0.9, -1.3, 0.6
    c0, c1, c2 = 0.2, -0.8, 1.1
    vel = lambda x: np.array([a1 + b11 * x[0] + b12 * x[1], a2 + b21 * x[0] - b11 * x[1]])
    grad = np.array([[b11, b12], [b21, -b11]])
    pres = lambda x: c0 + c1 * x[0] + c2 * x[1]
    sig = mu * grad

    def value(v, comp, x, n):
        if v in (0, 1):
            return vel(x)[v]
        if v == 2:
            return pres(x)
        if 3 <= v <= 6:
            return sig[(v - 3) // 2, (v - 3) % 2]
        if v in (7, 8):
            return vel(x)[v - 7]
        i = v - 9
        return pres(x) * n[i] - sig[i] @ n
    return value, (c1, c2)


@pytest.mark.parametrize("mu", [1.0, 0.5])
def test_stokes_vgp_linear_solution_has_zero_residual(mu):
    """Stokes VGP at the Stokes workload's orders (H1 order 4, test 5) on jittered
    quads with mixed parities: the exact linear solution gives l = B u (and a
    zero energy error); flipping the sign of any one of the 20 bilinear terms
    leaves a residual."""
    nodes, par = _quad_mesh(3, 3, 0.1, 8)
    par[:, 2] = -1
    f = oracle.form("stokes", 2, mu)
    bk = capi.BucketDesc(capi.QUAD, 4, 5, 0)
    value, grad_p = _stokes_solution(mu)
    load = _const_load(grad_p)
    u = nodal_trial_coefficients(f, 4, nodes, par, value)
    assert _residual(f, bk, nodes, par, u, load=load).max() < 1e-13
    eta2 = oracle.energy_error(f, bk, nodes, par, u, load=load)
    o = oracle.assemble(f, bk, nodes, par, load=load, want=("rhs",))
    assert np.abs(eta2).max() < 1e-24 * max(1.0, np.abs(o["rhs"]).max() ** 2)
    assert f.n_bterms == 20
    for t in range(f.n_bterms):
        g = oracle.form_like(f)
        g.bterms[t].coef.scale = -g.bterms[t].coef.scale
        assert _residual(g, bk, nodes, par, u, load=load).max() > 1e-4, t
    # mu enters as 1/mu in (sigma/mu, tau): the solution of the other mu fails
    wrong, _ = _stokes_solution(1.0 / mu if mu != 1.0 else 2.0)
    uw = nodal_trial_coefficients(f, 4, nodes, par, wrong)
    assert _residual(f, bk, nodes, par, uw, load=load).max() > 1e-4


def _cdr_case(seed):
    nodes, par = _quad_mesh(3, 3, 0.1, seed)
    par[:, 1] = -1
    C = nodes.shape[0]
    eps = 10.0 ** np.linspace(-3, 0, C)
    beta = 1.3 * np.array([np.cos(0.4), np.sin(0.4)])
    alpha = np.linspace(0.2, 1.0, C)
    return nodes, par, eps, beta, alpha


def test_cdr_constant_solution_pins_reaction_and_convective_flux():
    """ConvectionDiffusionReactionFormulation (ULTRAWEAK, 2D; terms of
    ConvectionDiffusionReactionFormulation.cpp:88-94) with per-cell alpha: u = c
    gives sigma = 0, sigma_n = -beta.n u * parity and f = alpha c. Each cell
    is solved with its own load. The residual vanishes; it does not when the
    reaction coefficient is doubled or the convective flux is dropped."""
    nodes, par, eps, beta, alpha = _cdr_case(12)
    f = oracle.form("cdr", 2)
    bk = capi.BucketDesc(capi.QUAD, 3, 4, 0)
    c0 = 0.75

    def coeffs(flux_scale):
        def value(v, comp, x, n):
            if v in (0, 2):
                return c0
            if v == 1:
                return 0.0
            return -flux_scale * (beta @ n) * c0
        return value

    for c in range(nodes.shape[0]):
        nd, pc = nodes[c:c + 1], par[c:c + 1]
        co = np.array([[eps[c], beta[0], beta[1], alpha[c]]])
        load = _const_load([alpha[c] * c0])
        u = nodal_trial_coefficients(f, 3, nd, pc, coeffs(1.0))
        assert _residual(f, bk, nd, pc, u, co, load)[0] < 1e-13, c
        co2 = co.copy()
        co2[0, 3] *= 2.0
        assert _residual(f, bk, nd, pc, u, co2, load)[0] > 1e-4, c
        u0 = nodal_trial_coefficients(f, 3, nd, pc, coeffs(0.0))
        assert _residual(f, bk, nd, pc, u0, co, load)[0] > 1e-4, c


def test_cdr_linear_solution_pins_sqrt_eps():
    """The same formulation with alpha = 0 and u linear: sigma = sqrt(eps) grad u
    (the scaled flux variable), sigma_n = (eps grad u - beta u).n * parity,
    f = beta . grad u. The residual vanishes for eps in [1e-3, 1]; with
    sigma = eps grad u (eps entering where sqrt(eps) belongs) it does not."""
    nodes, par, eps, beta, _ = _cdr_case(13)
    C = nodes.shape[0]
    f = oracle.form("cdr", 2)
    bk = capi.BucketDesc(capi.QUAD, 3, 4, 0)
    a = np.array([0.3, -0.9, 1.4])
    lin = lambda x: a[0] + a[1:] @ x
    coefs = np.stack([eps, np.full(C, beta[0]), np.full(C, beta[1]), np.zeros(C)], axis=1)
    load = _const_load([beta @ a[1:]])

    def interp(sig_of_eps):
        us = []
        for c in range(C):
            e = eps[c]

            def value(v, comp, x, n, e=e):
                if v in (0, 2):
                    return lin(x)
                if v == 1:
                    return sig_of_eps(e) * a[1 + comp]
                return (e * a[1:] @ n) - (beta @ n) * lin(x)
            us.append(nodal_trial_coefficients(f, 3, nodes[c:c + 1], par[c:c + 1], value)[0])
        return np.array(us)

    res = _residual(f, bk, nodes, par, interp(np.sqrt), coefs, load)
    assert res.max() < 1e-13
    wrong = _residual(f, bk, nodes, par, interp(lambda e: e), coefs, load)
    assert np.all(wrong[eps < 0.5] > 1e-5)


def test_hex_side_maps_follow_shards_face_nodes():
    """The shards Hexahedron<8> side-node table used by the tests: every side is
    a face of the reference cube, its nodes run around that face and the side
    parametrisation's normal (d xi/ds0 x d xi/ds1) points outward; and the
    library's trace/flux dof nodes on each side (dpg_skeleton_nodes) are the
    images of the side's GLL tensor nodes under that parametrisation
    (CamelliaCellToolsTests.cpp:154-230: MapToReferenceSubcell agrees with the
    node map)."""
    for s, face in enumerate(HEX_FACES):
        P = HEX_VERTS[list(face)]
        const = [d for d in range(3) if np.all(P[:, d] == P[0, d])]
        assert len(const) == 1
        for i in range(4):
            assert np.sum(P[i] != P[(i + 1) % 4]) == 1  # consecutive nodes share an edge
        n = np.cross(P[1] - P[0], P[3] - P[0])
        assert np.dot(n, P.mean(axis=0)) > 0, s
    assert sorted(tuple(sorted(fc)) for fc in HEX_FACES) == sorted(
        {tuple(sorted(np.nonzero(HEX_VERTS[:, d] == v)[0])) for d in range(3) for v in (-1, 1)})
    f = oracle.form("poisson", 3)
    k = 4
    bk = capi.BucketDesc(capi.HEX, k, 5, 0)
    xi = dpg.skeleton_nodes(f, bk)
    gt, gf = oracle.gll_nodes(k), oracle.gll_nodes(k - 1)
    nt, nfl = len(gt) ** 2, len(gf) ** 2
    want = np.concatenate([side_points(8, s, tensor_nodes(gt, 2)) for s in range(6)]
                          + [side_points(8, s, tensor_nodes(gf, 2)) for s in range(6)])
    assert xi.shape == (6 * (nt + nfl), 3)
    assert np.abs(xi - want).max() < 1e-14


def _sheared_hexes():
    nodes, par = dpg.mesh_hex(2, 2, 1, 0.0, 3)
    A = np.array([[1.0, 0.3, -0.2], [0.1, 0.9, 0.25], [-0.15, 0.2, 1.1]])
    return nodes @ A.T + np.array([0.1, -0.3, 0.2]), par


def test_hex_poisson_linear_solution_has_zero_residual():
    """3D Poisson ultraweak at the hex workload's orders (H1 order 4, test 5:
    N = 666, M = 502) on sheared affine hexes (planar faces, constant normals, so
    that sigma_n = grad u . n is a polynomial): u linear gives l = B u with
    f = 0. Flipping one face parity (or using a face node map that starts at
    another vertex) breaks it."""
    nodes, par = _sheared_hexes()
    f = oracle.form("poisson", 3)
    k = 4
    bk = capi.BucketDesc(capi.HEX, k, 5, 0)
    a = np.array([0.5, -0.7, 1.2, 0.4])
    lin = lambda x: a[0] + a[1:] @ x

    def value(v, comp, x, n):
        if v in (0, 2):
            return lin(x)
        if v == 1:
            return a[1 + comp]
        return a[1:] @ n

    u = nodal_trial_coefficients(f, k, nodes, par, value)
    assert u.shape[1] == oracle.sizes(f, bk)["M"] == 502
    assert _residual(f, bk, nodes, par, u).max() < 1e-13
    bad = par.copy()
    bad[0, 1] = -bad[0, 1]
    assert _residual(f, bk, nodes, bad, u)[0] > 1e-4
    # the same trace values placed with face 5 parametrised from its second vertex
    saved = HEX_FACES[5]
    try:
        HEX_FACES[5] = saved[1:] + saved[:1]
        ur = nodal_trial_coefficients(f, k, nodes, par, value)
    finally:
        HEX_FACES[5] = saved
    assert _residual(f, bk, nodes, par, ur).max() > 1e-4
