"""Seeded synthetic workloads shared by the benchmark and the tests.

hp convection-diffusion: the graded quad mesh of dpg_mesh_quad_graded with
triangles (the level-0 quads of the first coarse columns and every quad of
the first coarse rows, at all refinement levels, each split along its
diagonal) and a per-cell field degree p and test enrichment dp;
cells of equal (topology, p, dp) form one element type (bucket), as
Camellia's ElementType keys them (GlobalDofAssignment.cpp:615-644: trial
order p+1, test order p+dp)."""
import numpy as np

from . import capi, dpg

SEED_C3 = 0xC0FFEE + 3
SEED_C1 = 0xC0FFEE + 1


def split_quads(nodes, par):
    """Each quad (v0..v3) -> triangles (v0,v1,v2), (v0,v2,v3); the outer sides keep
    the quad's parities, the diagonal is +1 in the first triangle, -1 in the second."""
    C = nodes.shape[0]
    tri = np.zeros((2 * C, 3, 2))
    tp = np.zeros((2 * C, 3), dtype=np.int8)
    tri[0::2] = nodes[:, [0, 1, 2]]
    tri[1::2] = nodes[:, [0, 2, 3]]
    one = np.ones(C, np.int8)
    tp[0::2] = np.stack([par[:, 0], par[:, 1], one], axis=1)
    tp[1::2] = np.stack([-one, par[:, 2], par[:, 3]], axis=1)
    return tri, tp


def cd_hp(n0=16, levels=10, seed=SEED_C3, jitter=0.05, pmax=6, dpmax=3, tri_cols=4, tri_rows=1):
    """hp-adaptive 2D convection-diffusion: n0=16, levels=10, tri_cols=4,
    tri_rows=1 gives 46,215 quads and 6,290 triangles (52,505 cells, 12%
    triangles, 36 element types); tri_rows=0 gives 49,296 quads and 128
    triangles.

    Mesh graded toward the outflow edge x = 1; per-cell p ~ U{1..pmax} and
    dp ~ U{1..dpmax} (H1 order p+1, test degree p+dp); per-cell eps
    log-uniform on [1e-4, 1e-2]; beta one unit vector at a seeded angle for
    all cells; load: 8 sine modes with k in 1..6. The two triangles of a split
    quad inherit its p, dp and eps.
    Returns (form, cells, load) with cells[topology] = dict(nodes, parities,
    coefs, p, dp)."""
    f = dpg.form("cd", 2)
    nodes, par, lvl = dpg.mesh_quad_graded(n0, levels, jitter, seed)
    C = nodes.shape[0]
    p = np.minimum(np.floor(dpg.rng_uniform(seed + 1, 1.0, pmax + 1.0, C)), pmax).astype(np.int32)
    dp = np.minimum(np.floor(dpg.rng_uniform(seed + 2, 1.0, dpmax + 1.0, C)), dpmax).astype(np.int32)
    eps = 10.0 ** dpg.rng_uniform(seed + 3, -4.0, -2.0, C)
    theta = dpg.rng_uniform(seed + 4, 0.0, 2 * np.pi, 1)[0]
    coefs = np.stack([eps, np.full(C, np.cos(theta)), np.full(C, np.sin(theta))], axis=1)
    load = dpg.load_sine_modes(2, 1, 8, 6, SEED_C1 + 100)
    cx, cy = nodes[:, :, 0].mean(axis=1), nodes[:, :, 1].mean(axis=1)
    band = ((lvl == 0) & (cx < tri_cols / n0)) | (cy < tri_rows / n0)
    q = ~band
    cells = {capi.QUAD: dict(nodes=nodes[q], parities=par[q], coefs=coefs[q], p=p[q], dp=dp[q])}
    if band.any():
        tn, tp = split_quads(nodes[band], par[band])
        rep = lambda a: np.repeat(a[band], 2, axis=0)
        cells[capi.TRI] = dict(nodes=tn, parities=tp, coefs=rep(coefs), p=rep(p), dp=rep(dp))
    return f, cells, load


def hp_groups(cells):
    """Element types of an hp mesh, most work first: a list of dicts with the
    bucket (topology, h1_order, test_degree) and the type's cell arrays."""
    out = []
    for topo, c in cells.items():
        for a, b in sorted(set(zip(c["p"].tolist(), c["dp"].tolist()))):
            idx = np.nonzero((c["p"] == a) & (c["dp"] == b))[0]
            out.append(dict(topology=topo, k=a + 1, kt=a + b, bucket=dpg.bucket(topo, a + 1, a + b),
                            nodes=np.ascontiguousarray(c["nodes"][idx]),
                            parities=np.ascontiguousarray(c["parities"][idx]),
                            coefs=np.ascontiguousarray(c["coefs"][idx]), ncells=len(idx)))
    out.sort(key=lambda g: (-(g["kt"] + g["k"]), -g["k"], g["topology"]))
    return out
