"""Shared helpers: seeded workloads, error norms, exactly representable solutions."""
import numpy as np

from camellia2_amd import capi, dpg

# seeds: 0xC0FFEE + configuration number
SEED = {1: 0xC0FFEE + 1, 2: 0xC0FFEE + 2, 3: 0xC0FFEE + 3, 4: 0xC0FFEE + 4, 5: 0xC0FFEE + 5}


def rel_fro(a, b):
    """relative Frobenius error of a against the reference b, per cell, max over cells"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    axes = tuple(range(1, a.ndim))
    num = np.sqrt(np.sum((a - b) ** 2, axis=axes))
    den = np.sqrt(np.sum(b ** 2, axis=axes))
    return float(np.max(num / np.maximum(den, 1e-300)))


def poisson2d(nx, ny, k=3, kt=4, seed=SEED[1], jitter=0.10):
    """2D Poisson ultraweak, 8-mode sine load, jittered quads (configuration C1 at 16x16)."""
    f = dpg.form("poisson", 2)
    b = dpg.bucket(capi.QUAD, k, kt)
    nodes, par = dpg.mesh_quad(nx, ny, jitter, seed)
    load = dpg.load_sine_modes(2, 1, 8, 6, seed + 100)
    return f, b, nodes, par, None, load


def cdr2d(nx, ny, k=3, kt=4, seed=SEED[5], jitter=0.10):
    """2D convection-diffusion-reaction with per-cell [eps, beta_x, beta_y, c] (configuration C5 at 64x64)."""
    f = dpg.form("cdr", 2)
    b = dpg.bucket(capi.QUAD, k, kt)
    nodes, par = dpg.mesh_quad(nx, ny, jitter, seed)
    C = nx * ny
    eps = 10.0 ** dpg.rng_uniform(seed + 1, -3.0, 0.0, C)
    theta = dpg.rng_uniform(seed + 2, 0.0, 2 * np.pi, 1)[0]
    mag = dpg.rng_uniform(seed + 3, 0.5, 2.0, C)
    react = dpg.rng_uniform(seed + 4, 0.0, 1.0, C)
    coefs = np.stack([eps, mag * np.cos(theta), mag * np.sin(theta), react], axis=1)
    load = dpg.load_sine_modes(2, 1, 8, 6, SEED[1] + 100)
    return f, b, nodes, par, coefs, load


def stokes2d(nx, ny, k=4, kt=5, seed=SEED[2], jitter=0.10):
    """2D Stokes VGP ultraweak, 8 sine modes per force component (configuration C2 at 128x128)."""
    f = dpg.form("stokes", 2, 1.0)
    b = dpg.bucket(capi.QUAD, k, kt)
    nodes, par = dpg.mesh_quad(nx, ny, jitter, seed)
    load = dpg.load_sine_modes(2, 2, 8, 6, seed + 100)
    return f, b, nodes, par, None, load


def poisson3d(nx, ny, nz, k=4, kt=5, seed=SEED[4], jitter=0.05):
    """3D Poisson ultraweak, 16-mode sine load, jittered hexes (configuration C4 at 32^3)."""
    f = dpg.form("poisson", 3)
    b = dpg.bucket(capi.HEX, k, kt)
    nodes, par = dpg.mesh_hex(nx, ny, nz, jitter, seed)
    load = dpg.load_sine_modes(3, 1, 16, 6, seed + 100)
    return f, b, nodes, par, None, load



# the hp workload lives in the package (bench.py uses it too)
from camellia2_amd.workloads import cd_hp, hp_groups  # noqa: E402,F401


REF_VERTS = {4: np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]]),
             3: np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])}


def _vertex_basis(ref, xi):
    if len(ref) == 3:
        return np.array([1.0 - xi[0] - xi[1], xi[0], xi[1]])
    return np.array([(1 + ref[v, 0] * xi[0]) * (1 + ref[v, 1] * xi[1]) / 4 for v in range(4)])


def linear_trial_coefficients(nodes, par, k, a, eps=None, beta=None):
    """Trial coefficients of u = a0 + a1 x + a2 y on quads (C,4,2) or triangles
    (C,3,2) by nodal interpolation: the field at its nodes (GLL tensor nodes on
    quads, warp-and-blend nodes on triangles, degree k-1), traces (Cn(k)) and
    fluxes (Cn(k-1)) at the GLL nodes of each side.
    Poisson UW (beta None): sigma = grad u, uhat = u, sigma_n_hat = sigma.n * parity.
    ConvectionDiffusion (per-cell eps, constant beta): sigma = eps grad u,
    uhat = u, tc = (beta u - sigma).n * parity."""
    from oracle import oracle
    C, V = nodes.shape[0], nodes.shape[1]
    ref = REF_VERTS[V]
    gf, gt = oracle.gll_nodes(k - 1), oracle.gll_nodes(k)
    if V == 3:
        fnodes = oracle.tri_nodes(k - 1)
    else:
        fnodes = np.array([(gf[i], gf[j]) for j in range(len(gf)) for i in range(len(gf))])
    nfld, nt, nfl = len(fnodes), len(gt), len(gf)
    M = 3 * nfld + V * nt + V * nfl
    u = np.zeros((C, M))
    sides = [(s, (s + 1) % V) for s in range(V)]

    def phys(c, xi):
        return _vertex_basis(ref, xi) @ nodes[c]

    def lin(x):
        return a[0] + a[1] * x[0] + a[2] * x[1]

    for c in range(C):
        e = 1.0 if beta is None else eps[c]
        u[c, :nfld] = [lin(phys(c, p)) for p in fnodes]
        u[c, nfld:2 * nfld] = e * a[1]
        u[c, 2 * nfld:3 * nfld] = e * a[2]
        off = 3 * nfld
        for s, (p, q) in enumerate(sides):
            for i, t in enumerate(gt):
                u[c, off + s * nt + i] = lin(phys(c, 0.5 * (ref[p] + ref[q]) + 0.5 * (ref[q] - ref[p]) * t))
        off += V * nt
        for s, (p, q) in enumerate(sides):
            T = nodes[c, q] - nodes[c, p]
            n = np.array([T[1], -T[0]]) / np.hypot(*T)
            dudn = a[1] * n[0] + a[2] * n[1]
            for i, t in enumerate(gf):
                if beta is None:
                    val = dudn
                else:
                    x = phys(c, 0.5 * (ref[p] + ref[q]) + 0.5 * (ref[q] - ref[p]) * t)
                    val = (beta[0] * n[0] + beta[1] * n[1]) * lin(x) - e * dudn
                u[c, off + s * nfl + i] = val * par[c, s]
    return u


from camellia2_amd.workloads import split_quads  # noqa: E402,F401


# shards Hexahedron<8> (the reference's hex topology, CamelliaCellTools.cpp:1043-1122
# parametrises a side through it): vertices and the side-node map, transcribed
# from the shards cell-topology tables (shards is part of Trilinos and not
# vendored in the reference). Face s is the image of the reference quad whose
# vertex i maps to hex vertex HEX_FACES[s][i].
HEX_VERTS = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
HEX_FACES = [(0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 4, 7, 3), (0, 3, 2, 1), (4, 5, 6, 7)]


def ref_vertices(nverts):
    return HEX_VERTS if nverts == 8 else REF_VERTS[nverts]


def side_points(nverts, side, t):
    """reference (volume) coordinates of side points with side coordinates t (n, dim-1):
    the affine map taking the reference line/quad vertices to the side's vertices"""
    t = np.atleast_2d(t)
    if nverts == 8:
        n = HEX_VERTS[list(HEX_FACES[side])]
        c0 = n.mean(axis=0)
        c1 = 0.25 * (-n[0] + n[1] + n[2] - n[3])
        c2 = 0.25 * (-n[0] - n[1] + n[2] + n[3])
        return c0 + np.outer(t[:, 0], c1) + np.outer(t[:, 1], c2)
    ref = REF_VERTS[nverts]
    a, b = ref[side], ref[(side + 1) % nverts]
    return 0.5 * (a + b) + np.outer(t[:, 0], 0.5 * (b - a))


def vertex_map(nodes_c, xi):
    """physical points of reference points xi (n, dim) under the C1 vertex map of one cell"""
    V, dim = nodes_c.shape
    ref = ref_vertices(V)
    out = []
    for p in np.atleast_2d(xi):
        if V == 3:
            w = np.array([1.0 - p[0] - p[1], p[0], p[1]])
        else:
            w = np.prod([(1 + ref[:, d] * p[d]) / 2 for d in range(dim)], axis=0)
        out.append(w @ nodes_c)
    return np.array(out)


def tensor_nodes(g, dim):
    """tensor points of the 1D nodes g, x fastest: (len(g)^dim, dim)"""
    grids = np.meshgrid(*([g] * dim), indexing="ij")
    return np.stack([gr.ravel(order="F") for gr in grids], axis=1) if dim > 1 else g[:, None]


def nodal_trial_coefficients(f, k, nodes, par, value):
    """Trial coefficients (C, M) of a quad/hex element of H1 order k by nodal
    interpolation in the reference's dof layout (variables in order; fields at
    the GLL tensor nodes of degree k-1, vector fields component-major; traces
    (Cn(k)) and fluxes (Cn(k-1)) at the GLL nodes of every side, side after side).
    value(var, comp, x, n) gives the variable's value at the physical point x
    (n: the outward unit normal on sides, None in the volume); flux values are
    multiplied by the side parity (LinearTerm.cpp:518-524)."""
    from oracle import oracle
    C, V, dim = nodes.shape
    S = 2 * dim if V != 3 else 3
    gf, gt = oracle.gll_nodes(k - 1), oracle.gll_nodes(k)
    cols = []
    for c in range(C):
        xc = nodes[c]
        cen = xc.mean(axis=0)
        row = []
        for v in range(f.n_trial):
            kind = f.trial_kind[v]
            if kind in (capi.FIELD, capi.FIELD_VECTOR):
                X = vertex_map(xc, tensor_nodes(gf, dim))
                for comp in range(dim if kind == capi.FIELD_VECTOR else 1):
                    row += [value(v, comp, x, None) for x in X]
                continue
            g = gt if kind == capi.TRACE else gf
            tn = tensor_nodes(g, dim - 1)
            for s in range(S):
                X = vertex_map(xc, side_points(V, s, tn))
                # outward unit normal of the (planar) side
                P = vertex_map(xc, side_points(V, s, np.zeros((1, dim - 1))))[0]  # side centre
                if dim == 2:
                    a, b = vertex_map(xc, side_points(V, s, np.array([[-1.0], [1.0]])))
                    T = b - a
                    n = np.array([T[1], -T[0]])
                else:
                    q = vertex_map(xc, side_points(V, s, np.array([[-1.0, -1.0], [1.0, -1.0], [-1.0, 1.0]])))
                    n = np.cross(q[1] - q[0], q[2] - q[0])
                n = n / np.linalg.norm(n)
                if np.dot(n, P - cen) < 0:
                    n = -n
                sgn = par[c, s] if kind == capi.FLUX else 1.0
                row += [sgn * value(v, 0, x, n) for x in X]
        cols.append(row)
    return np.array(cols)


def hp_mesh(n0, levels, tri_cols, pmax=3, dpmax=2):
    """cells of the graded hp mesh in one list (quads, then triangles): nodes
    per topology, the bucket of every cell, the buckets, global vertex ids and
    the vertices' coordinates"""
    from camellia2_amd.workloads import cd_hp

    _, cells, _ = cd_hp(n0, levels, tri_cols=tri_cols, pmax=pmax, dpmax=dpmax, tri_rows=0)
    order = []  # (topology, index within the topology)
    types = {}
    for topo in (capi.QUAD, capi.TRI):
        if topo not in cells:
            continue
        c = cells[topo]
        for i in range(c["nodes"].shape[0]):
            key = (topo, int(c["p"][i]) + 1, int(c["p"][i] + c["dp"][i]))
            types.setdefault(key, len(types))
            order.append((topo, i, types[key]))
    buckets = [None] * len(types)
    for (topo, k, kt), b in types.items():
        buckets[b] = dpg.bucket(topo, k, kt)
    # global vertex ids from the coordinates (shared vertices are identical)
    pts = np.concatenate([cells[t]["nodes"].reshape(-1, 2) for t in (capi.QUAD, capi.TRI) if t in cells])
    X, inv = np.unique(pts, axis=0, return_inverse=True)
    cv = np.full((len(order), 4), -1, dtype=np.int32)
    pos, row = 0, 0
    for t in (capi.QUAD, capi.TRI):
        if t not in cells:
            continue
        V = cells[t]["nodes"].shape[1]
        n = cells[t]["nodes"].shape[0]
        cv[row:row + n, :V] = inv[pos:pos + n * V].reshape(n, V)
        pos += n * V
        row += n
    cell_bucket = np.array([o[2] for o in order], dtype=np.int32)
    return cells, order, buckets, cell_bucket, cv, X


def per_bucket(cells, order, nb):
    """the nodes, parities and global cell ids of every bucket's cells, in cell order"""
    out = []
    for b in range(nb):
        ids = [g for g, o in enumerate(order) if o[2] == b]
        nodes = np.stack([cells[order[g][0]]["nodes"][order[g][1]] for g in ids])
        par = np.stack([cells[order[g][0]]["parities"][order[g][1]] for g in ids])
        out.append(dict(ids=ids, nodes=nodes, parities=par))
    return out
