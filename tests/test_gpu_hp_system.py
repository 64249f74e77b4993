"""The condensed system of an hp mesh (dpg_hp_system): the graded quad mesh
with hanging sides, a band of triangles and a polynomial order and test
enrichment per cell, through local assembly of every element type, the
minimum-rule constrained assembly, the global solve, the local trace and flux
coefficients and the field recovery. A linear exact solution prescribed on the
boundary comes back in every trial variable of every cell (hanging sides and
order jumps included), and on a conforming single-type mesh the hp system
solves to the same local coefficients as the conforming skeleton."""
import numpy as np
import pytest

from camellia2_amd import capi, dpg

from helpers import hp_mesh, linear_trial_coefficients, per_bucket

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("shape", [(4, 3, 1), (6, 4, 2)])
def test_linear_solution_on_an_hp_mesh_with_hanging_sides(shape):
    n0, levels, tri_cols = shape
    f = dpg.form("poisson", 2)
    cells, order, buckets, cell_bucket, cv, X = hp_mesh(n0, levels, tri_cols)
    nb = len(buckets)
    groups = per_bucket(cells, order, nb)
    A = dpg.LocalAssembler(f, buckets)
    outs = []
    for b, g in enumerate(groups):
        o, nbad = A.local_stiffness_and_rhs(b, dev(g["nodes"]), dev(g["parities"]), None, None, want_schur=True)
        assert nbad == 0
        outs.append(o)
    hs = dpg.HpSystem(f, buckets, cell_bucket, cv, X)
    vals, rhs = hs.assemble([o["schur"] for o in outs], [o["schur_rhs"] for o in outs])
    a = np.array([0.4, -0.9, 1.3])
    exact = [linear_trial_coefficients(g["nodes"], g["parities"], buckets[b].h1_order, a)
             for b, g in enumerate(groups)]
    fixed, entity, rep_cell, rep_local = hs.dofs()
    x0 = np.zeros(hs.nglobal)
    where = {gid: (b, j) for b, g in enumerate(groups) for j, gid in enumerate(g["ids"])}
    sz = [dpg.sizes(f, bk) for bk in buckets]
    for d in np.flatnonzero(fixed):
        assert rep_cell[d] >= 0
        b, j = where[rep_cell[d]]
        x0[d] = exact[b][j, sz[b].F + rep_local[d]]
    x = dev(x0)
    it, rel = hs.solve(vals, rhs, dev(fixed), x, tol=1e-13, maxit=20000)
    print(f"{len(order)} cells ({nb} element types), {hs.nglobal} global dofs: CG {it} iterations, relres {rel:.1e}")
    assert rel <= 1e-13
    u = [torch.zeros((len(g["ids"]), sz[b].M), dtype=torch.float64, device="cuda:0") for b, g in enumerate(groups)]
    hs.local(x, u)
    scale = max(np.abs(e).max() for e in exact)
    for b in range(nb):
        assert A.recover_fields(b, outs[b]["K"], outs[b]["rhs"], u[b]) == 0
        err = np.abs(u[b].cpu().numpy() - exact[b]).max()
        # the trace solve stops at a relative residual of 1e-13; the graded
        # mesh's global matrix (hanging sides down to h0 / 2^(levels-1)) carries
        # it into the traces and the recovered fields at up to ~1e5 times that
        assert err <= 2e-8 * scale, (b, buckets[b].topology, buckets[b].h1_order, err)
        eta2 = A.energy_error(b, dev(groups[b]["nodes"]), dev(groups[b]["parities"]), u[b], None, None)
        assert eta2.max().item() <= 1e-16
    A.close()
    hs.close()


def test_hp_system_equals_the_conforming_skeleton_on_a_uniform_mesh():
    f = dpg.form("poisson", 2)
    b = dpg.bucket(capi.QUAD, 3, 4)
    nodes, par = dpg.mesh_quad(7, 5, 0.1, 211)
    load = dpg.load_sine_modes(2, 1, 4, 3, 0xC0FFEE + 111)
    A = dpg.LocalAssembler(f, [b])
    o, nbad = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), None, load, want_schur=True)
    assert nbad == 0
    cv = dpg.vertex_ids(nodes)
    s = dpg.sizes(f, b)
    C = nodes.shape[0]
    # conforming skeleton
    gdof, n = dpg.skeleton_numbering(f, b, cv)
    sysm = dpg.CondensedSystem(gdof, n)
    v1, r1 = sysm.assemble(o["schur"], o["schur_rhs"])
    fx1 = dpg.boundary_trace_mask(f, b, gdof, cv, n)
    x1 = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    it1, rel1 = sysm.solve(v1, r1, dev(fx1), x1, tol=1e-13, maxit=20000)
    u1 = x1[torch.from_numpy(gdof.astype(np.int64)).to("cuda:0")].cpu().numpy()
    sysm.close()
    # hp system
    X = np.zeros((cv.max() + 1, 2))
    X[cv.ravel()] = nodes.reshape(-1, 2)
    hs = dpg.HpSystem(f, [b], np.zeros(C, dtype=np.int32), cv, X)
    assert hs.nglobal == n
    v2, r2 = hs.assemble([o["schur"]], [o["schur_rhs"]])
    fx2, _, _, _ = hs.dofs()
    assert fx2.sum() == fx1.sum()
    x2 = torch.zeros(hs.nglobal, dtype=torch.float64, device="cuda:0")
    it2, rel2 = hs.solve(v2, r2, dev(fx2), x2, tol=1e-13, maxit=20000, blocks=False)
    u2 = torch.zeros((C, s.M), dtype=torch.float64, device="cuda:0")
    hs.local(x2, [u2])
    u2 = u2[:, s.F:].cpu().numpy()
    print(f"conforming: {it1} iterations, hp system: {it2} iterations")
    assert rel1 <= 1e-13 and rel2 <= 1e-13
    assert np.abs(u2 - u1).max() <= 1e-10 * np.abs(u1).max()
    hs.close()
    A.close()
