"""The solve loop on the device after local assembly: global condensed system,
CG solve with the trace prescribed on the boundary, field recovery per cell,
energy error per cell and marking -- checked end to end: a linear exact
solution comes back in every trial variable with zero energy error, the CG
solution matches a direct solve of the same system, and for a sine load the
energy error falls under refinement."""
import numpy as np
import pytest

from camellia2_amd import capi, dpg

from helpers import linear_trial_coefficients

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
sp = pytest.importorskip("scipy.sparse")
spla = pytest.importorskip("scipy.sparse.linalg")


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def boundary_trace_mask(f, b, gdof, cv, n):
    """1 on the global trace dofs of boundary sides (sides of one cell only)"""
    V = cv.shape[1]
    sides = [(i, (i + 1) % V) for i in range(V)]
    nt = b.h1_order + 1
    count = {}
    for c in range(cv.shape[0]):
        for s in sides:
            key = tuple(sorted(cv[c, list(s)]))
            count[key] = count.get(key, 0) + 1
    mask = np.zeros(n, dtype=np.int32)
    for c in range(cv.shape[0]):
        for si, s in enumerate(sides):
            if count[tuple(sorted(cv[c, list(s)]))] == 1:
                mask[gdof[c, si * nt:(si + 1) * nt]] = 1
    return mask


def solve_loop(f, b, nodes, par, load, fixed_values=None, tol=1e-13):
    """local assembly -> condensed global system -> CG -> field recovery ->
    energy error; returns (u (C,M) device, eta2 (C) device, iterations, system pieces)"""
    A = dpg.LocalAssembler(f, [b])
    dn, dp = dev(nodes), dev(par)
    out, nbad = A.local_stiffness_and_rhs(0, dn, dp, None, load, want_schur=True)
    assert nbad == 0
    cv = dpg.vertex_ids(nodes)
    gdof, n = dpg.skeleton_numbering(f, b, cv)
    sysm = dpg.CondensedSystem(gdof, n)
    vals, rhs = sysm.assemble(out["schur"], out["schur_rhs"])
    fixed = boundary_trace_mask(f, b, gdof, cv, n)
    x = np.zeros(n) if fixed_values is None else np.where(fixed == 1, fixed_values, 0.0)
    dx, dfixed = dev(x), dev(fixed)
    iters, rel = sysm.solve(vals, rhs, dfixed, dx, tol=tol, maxit=20000)
    s = dpg.sizes(f, b)
    u = torch.zeros((nodes.shape[0], s.M), dtype=torch.float64, device="cuda:0")
    u[:, s.F:] = dx[torch.from_numpy(gdof.astype(np.int64)).to("cuda:0")]
    assert A.recover_fields(0, out["K"], out["rhs"], u) == 0
    eta2 = A.energy_error(0, dn, dp, u, None, load)
    torch.cuda.synchronize()
    A.close()
    pieces = dict(sysm=sysm, vals=vals, rhs=rhs, fixed=fixed, x=dx, gdof=gdof, n=n)
    return u, eta2, iters, rel, pieces


def test_linear_solution_through_the_whole_loop():
    f = dpg.form("poisson", 2)
    b = dpg.bucket(capi.QUAD, 3, 5)
    nodes, par = dpg.mesh_quad(7, 6, 0.15, 101)
    a = np.array([0.3, -1.2, 0.7])
    exact = linear_trial_coefficients(nodes, par, 3, a)
    s = dpg.sizes(f, b)
    cv = dpg.vertex_ids(nodes)
    gdof, n = dpg.skeleton_numbering(f, b, cv)
    xg = np.zeros(n)
    xg[gdof.ravel()] = exact[:, s.F:].ravel()
    u, eta2, iters, rel, pc = solve_loop(f, b, nodes, par, None, fixed_values=xg)
    print(f"{n} global dofs, CG {iters} iterations, relres {rel:.1e}")
    assert rel <= 1e-13
    x = pc["x"].cpu().numpy()
    assert np.abs(x - xg).max() <= 1e-9 * np.abs(xg).max()
    un = u.cpu().numpy()
    assert np.abs(un - exact).max() <= 1e-9 * np.abs(exact).max()  # fields recovered too
    assert eta2.max().item() <= 1e-18
    # the CG solution against a direct solve of the same system
    rp, col = pc["sysm"].pattern()
    Ah = sp.csr_matrix((pc["vals"].cpu().numpy(), col.cpu().numpy(), rp.cpu().numpy()), shape=(n, n))
    free = np.flatnonzero(pc["fixed"] == 0)
    fx = np.flatnonzero(pc["fixed"] == 1)
    xd = xg.copy()
    xd[free] = spla.spsolve(Ah[free][:, free].tocsc(), pc["rhs"].cpu().numpy()[free] - Ah[free][:, fx] @ xg[fx])
    assert np.abs(x - xd).max() <= 1e-10 * np.abs(xd).max()
    pc["sysm"].close()


def test_energy_error_falls_under_refinement_and_marking():
    f = dpg.form("poisson", 2)
    b = dpg.bucket(capi.QUAD, 3, 5)
    load = dpg.load_sine_modes(2, 1, 4, 3, 0xC0FFEE + 111)  # modes vanish on the boundary: u = 0 there
    totals = []
    for nx in (4, 8):
        nodes, par = dpg.mesh_quad(nx, nx, 0.1, 121)
        u, eta2, iters, rel, pc = solve_loop(f, b, nodes, par, load)
        assert rel <= 1e-12
        pc["sysm"].close()
        err = torch.sqrt(eta2).contiguous()
        totals.append(float(torch.sqrt(eta2.sum())))
        flags, cnt = dpg.mark_cells(err, 0.5)
        e = err.cpu().numpy()
        want = e > 0.5 * e.max()
        assert cnt == int(want.sum()) and np.array_equal(flags.cpu().numpy() == 1, want)
        print(f"{nx}x{nx}: CG {iters} iterations, energy error {totals[-1]:.3e}, {cnt} cells marked")
    assert totals[1] < totals[0] / 4.0


def test_riesz_representation_solves_the_gram_system():
    """e = G^-1 (l - B u) per cell against a direct solve with the oracle's G, B,
    l on the same cells, and r^T e against the energy error."""
    from oracle import oracle

    from helpers import cdr2d, rel_fro

    f, b, nodes, par, coefs, load = cdr2d(4, 4)
    s = dpg.sizes(f, b)
    u = np.asarray(dpg.rng_normal(17, nodes.shape[0] * s.M)).reshape(nodes.shape[0], s.M)
    A = dpg.LocalAssembler(f, [b])
    args = (dev(nodes), dev(par), dev(u))
    e = A.riesz_representation(0, args[0], args[1], args[2], dev(coefs), load).cpu().numpy()
    eta2 = A.energy_error(0, args[0], args[1], args[2], dev(coefs), load).cpu().numpy()
    A.close()
    o = oracle.assemble(oracle.form_like(f), b, nodes, par, coefs, load, want=("B", "G", "l"))
    r = o["l"] - np.einsum("cmn,cm->cn", o["B"], u)
    ex = np.linalg.solve(o["G"], r[..., None])[..., 0]
    err = rel_fro(e, ex)
    print(f"riesz representation vs direct solve: {err:.2e}")
    assert err < 1e-9
    assert np.allclose(np.einsum("cn,cn->c", r, e), eta2, rtol=1e-9, atol=0)


@pytest.mark.parametrize("dim", [2, 3])
def test_block_jacobi_solve_equals_direct_and_needs_fewer_iterations(dim):
    """Block-Jacobi over mesh entities (a side's fluxes with the traces interior
    to it, edge and vertex traces): the CG solution equals a direct solve of
    the same system and takes fewer iterations than point Jacobi."""
    if dim == 2:
        f, b = dpg.form("poisson", 2), dpg.bucket(capi.QUAD, 3, 5)
        nodes, par = dpg.mesh_quad(10, 9, 0.15, 131)
    else:
        f, b = dpg.form("poisson", 3), dpg.bucket(capi.HEX, 2, 3)
        nodes, par = dpg.mesh_hex(4, 3, 3, 0.1, 141)
    load = dpg.load_sine_modes(dim, 1, 4, 3, 0xC0FFEE + 151)
    A = dpg.LocalAssembler(f, [b])
    out, nbad = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), None, load, want_schur=True)
    A.close()
    assert nbad == 0
    cv = dpg.vertex_ids(nodes)
    gdof, n = dpg.skeleton_numbering(f, b, cv)
    sysm = dpg.CondensedSystem(gdof, n)
    vals, rhs = sysm.assemble(out["schur"], out["schur_rhs"])
    fixed = dpg.boundary_trace_mask(f, b, gdof, cv, n)
    dfixed = dev(fixed)
    x1 = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    it1, rel1 = sysm.solve(vals, rhs, dfixed, x1, tol=1e-12, maxit=20000)
    ent, ne = dpg.skeleton_entities(f, b, cv, n)
    sysm.set_blocks(ent, ne)
    x2 = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    it2, rel2 = sysm.solve(vals, rhs, dfixed, x2, tol=1e-12, maxit=20000)
    sysm.set_blocks(ent, 0)  # back to point Jacobi: the same iterations as before
    x3 = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    it3, _ = sysm.solve(vals, rhs, dfixed, x3, tol=1e-12, maxit=20000)
    rp, col = sysm.pattern()
    Ah = sp.csr_matrix((vals.cpu().numpy(), col.cpu().numpy(), rp.cpu().numpy()), shape=(n, n))
    free = np.flatnonzero(fixed == 0)
    xd = np.zeros(n)
    xd[free] = spla.spsolve(Ah[free][:, free].tocsc(), rhs.cpu().numpy()[free])
    sysm.close()
    e1 = np.abs(x1.cpu().numpy() - xd).max() / np.abs(xd).max()
    e2 = np.abs(x2.cpu().numpy() - xd).max() / np.abs(xd).max()
    print(f"dim {dim}: {n} dofs, {ne} entities; Jacobi {it1} iterations ({e1:.1e}), block-Jacobi {it2} ({e2:.1e})")
    assert rel1 <= 1e-12 and rel2 <= 1e-12
    assert e1 < 1e-9 and e2 < 1e-9
    assert it2 < it1 and it3 == it1
    assert torch.equal(x3, x1)


@pytest.mark.parametrize("dim", [2, 3])
def test_two_level_solve_equals_direct_and_needs_fewer_iterations(dim):
    """Two levels (entity block-Jacobi smoothing + a correction on the
    lowest-order trace/flux space of the same mesh): the flexible-CG solution
    equals a direct solve, and it takes fewer iterations than block-Jacobi CG;
    a linear exact solution prescribed on the boundary comes back exactly."""
    if dim == 2:
        f, b = dpg.form("poisson", 2), dpg.bucket(capi.QUAD, 4, 5)
        nodes, par = dpg.mesh_quad(16, 14, 0.15, 161)
    else:
        f, b = dpg.form("poisson", 3), dpg.bucket(capi.HEX, 3, 4)
        nodes, par = dpg.mesh_hex(5, 4, 4, 0.1, 171)
    load = dpg.load_sine_modes(dim, 1, 4, 3, 0xC0FFEE + 181)
    A = dpg.LocalAssembler(f, [b])
    out, nbad = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), None, load, want_schur=True)
    A.close()
    assert nbad == 0
    cv = dpg.vertex_ids(nodes)
    gdof, n = dpg.skeleton_numbering(f, b, cv)
    sysm = dpg.CondensedSystem(gdof, n)
    vals, rhs = sysm.assemble(out["schur"], out["schur_rhs"])
    fixed = dpg.boundary_trace_mask(f, b, gdof, cv, n)
    dfixed = dev(fixed)
    ent, ne = dpg.skeleton_entities(f, b, cv, n)
    sysm.set_blocks(ent, ne)
    co = dpg.CoarseLevel(f, b, cv)
    x1 = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    it1, rel1 = sysm.solve(vals, rhs, dfixed, x1, tol=1e-12, maxit=20000)
    x2 = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    it2, rel2 = sysm.solve_twolevel(co, out["schur"], vals, rhs, dfixed, x2, tol=1e-12, maxit=20000)
    rp, col = sysm.pattern()
    Ah = sp.csr_matrix((vals.cpu().numpy(), col.cpu().numpy(), rp.cpu().numpy()), shape=(n, n))
    free = np.flatnonzero(fixed == 0)
    xd = np.zeros(n)
    xd[free] = spla.spsolve(Ah[free][:, free].tocsc(), rhs.cpu().numpy()[free])
    e1 = np.abs(x1.cpu().numpy() - xd).max() / np.abs(xd).max()
    e2 = np.abs(x2.cpu().numpy() - xd).max() / np.abs(xd).max()
    print(f"dim {dim}: {n} dofs, coarse {co.nglobal} dofs / {co.nnz} nonzeros; block-Jacobi CG {it1} iterations "
          f"({e1:.1e}), two-level {it2} ({e2:.1e})")
    assert rel1 <= 1e-12 and rel2 <= 1e-12
    assert e1 < 1e-9 and e2 < 1e-9
    assert it2 < it1
    co.close()
    sysm.close()


def test_two_level_solve_returns_a_linear_exact_solution():
    f = dpg.form("poisson", 2)
    b = dpg.bucket(capi.QUAD, 3, 5)
    nodes, par = dpg.mesh_quad(9, 7, 0.15, 191)
    exact = linear_trial_coefficients(nodes, par, 3, np.array([0.3, -1.2, 0.7]))
    s = dpg.sizes(f, b)
    A = dpg.LocalAssembler(f, [b])
    out, nbad = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), None, None, want_schur=True)
    A.close()
    assert nbad == 0
    cv = dpg.vertex_ids(nodes)
    gdof, n = dpg.skeleton_numbering(f, b, cv)
    xg = np.zeros(n)
    xg[gdof.ravel()] = exact[:, s.F:].ravel()
    sysm = dpg.CondensedSystem(gdof, n)
    vals, rhs = sysm.assemble(out["schur"], out["schur_rhs"])
    fixed = dpg.boundary_trace_mask(f, b, gdof, cv, n)
    x = dev(np.where(fixed == 1, xg, 0.0))
    co = dpg.CoarseLevel(f, b, cv)
    it, rel = sysm.solve_twolevel(co, out["schur"], vals, rhs, dev(fixed), x, tol=1e-13, maxit=5000)
    print(f"two-level: {it} iterations, relres {rel:.1e}")
    assert rel <= 1e-13
    assert np.abs(x.cpu().numpy() - xg).max() <= 1e-9 * np.abs(xg).max()
    co.close()
    sysm.close()
