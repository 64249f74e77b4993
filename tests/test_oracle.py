"""Pins the oracle (CPU restatement of the reference's local assembly) against
the reference's own golden vectors and against what the mathematics fixes."""
import json
import os

import numpy as np
import pytest

from camellia2_amd import capi
from oracle import oracle

from helpers import linear_trial_coefficients, split_quads


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
