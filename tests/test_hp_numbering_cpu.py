"""The minimum-rule numbering of an hp mesh (dpg_hp_numbering, host) checked
with the oracle on the CPU: on the graded quad mesh with hanging sides, a
triangle band and per-cell orders, the local trace and flux coefficients of a
linear exact solution are exactly the constrained image M x of one global
vector (the constraints are conforming across hanging sides and order jumps),
and that vector satisfies the assembled condensed system sum_c M_c^T S_c M_c
built from the oracle's Schur complements."""
import numpy as np
import pytest

from camellia2_amd import capi, dpg
from oracle import oracle

from helpers import hp_mesh, linear_trial_coefficients, per_bucket

sp = pytest.importorskip("scipy.sparse")
spla = pytest.importorskip("scipy.sparse.linalg")


@pytest.mark.parametrize("shape", [(3, 2, 1), (4, 3, 1)])
def test_linear_solution_is_in_the_constrained_space_and_solves_the_system(shape):
    f = dpg.form("poisson", 2)
    cells, order, buckets, cell_bucket, cv, X = hp_mesh(*shape)
    nb = len(buckets)
    n, off, dofs, rp, ed, ew = dpg.hp_numbering(f, buckets, cell_bucket, cv, X)
    C = len(order)
    sz = [dpg.sizes(f, b) for b in buckets]
    groups = per_bucket(cells, order, nb)
    where = {gid: (b, j) for b, g in enumerate(groups) for j, gid in enumerate(g["ids"])}
    a = np.array([0.4, -0.9, 1.3])
    exact = [linear_trial_coefficients(g["nodes"], g["parities"], buckets[b].h1_order, a)
             for b, g in enumerate(groups)]
    ff = oracle.form_like(f)
    schur = []
    for b, g in enumerate(groups):
        o = oracle.assemble(ff, buckets[b], g["nodes"], g["parities"], None, None, want=("S", "g"))
        assert np.all(o["info"] == 0)
        schur.append(o)
    # M over all local trace/flux rows (cells in order) and the exact local values
    Mrows = sp.csr_matrix((ew, ed, rp), shape=(rp.shape[0] - 1, n))
    u_loc = np.concatenate([exact[where[c][0]][where[c][1], sz[where[c][0]].F:] for c in range(C)])
    x, istop = spla.lsqr(Mrows, u_loc, atol=1e-15, btol=1e-15, iter_lim=20000)[:2]
    e_rep = np.abs(Mrows @ x - u_loc).max() / np.abs(u_loc).max()
    # the assembled system sum_c M_c^T S_c M_c, from the oracle's Schur blocks
    rows, cols, vals = [], [], []
    rhs = np.zeros(n)
    r0 = 0
    for c in range(C):
        b, j = where[c]
        T = sz[b].T
        Mc = Mrows[r0:r0 + T]
        r0 += T
        Ac = (Mc.T @ sp.csr_matrix(schur[b]["S"][j]) @ Mc).tocoo()
        rows.append(Ac.row)
        cols.append(Ac.col)
        vals.append(Ac.data)
        rhs += Mc.T @ schur[b]["g"][j]
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    res = np.abs(A @ x - rhs).max() / (abs(A).max() * np.abs(x).max())
    print(f"{C} cells, {nb} element types, {n} global dofs: |Mx - u_local| {e_rep:.1e}, |Ax - b| {res:.1e}")
    assert e_rep < 1e-12
    assert res < 1e-12
    # a sanity control: the exact values of one cell changed leave the constrained space
    u_bad = u_loc.copy()
    u_bad[3] += 1e-3
    xb = spla.lsqr(Mrows, u_bad, atol=1e-15, btol=1e-15, iter_lim=20000)[0]
    assert np.abs(Mrows @ xb - u_bad).max() > 1e-5
