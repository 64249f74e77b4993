"""Global numbering of the condensed (trace and flux) dofs of conforming meshes
(dpg_skeleton_numbering), checked on the CPU against what the numbering must
mean geometrically: two local dofs share a global index exactly when they are
the same H^1 trace at the same physical node, or the same L^2 flux on the same
side at the same physical node. Random relabellings of the vertex ids put
every edge and face of the mesh in every orientation the canonical rule knows."""
import numpy as np
import pytest

from camellia2_amd import capi, dpg

from helpers import split_quads

HEX_REF = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)
QUAD_REF = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=float)


def vertex_basis(topo, xi):
    """(P, V) values of the cell's vertex basis at reference points xi (P, dim)"""
    if topo == capi.TRI:
        return np.stack([1 - xi[:, 0] - xi[:, 1], xi[:, 0], xi[:, 1]], axis=1)
    ref = QUAD_REF if topo == capi.QUAD else HEX_REF
    return np.prod(1 + xi[:, None, :] * ref[None, :, :], axis=2) / 2 ** ref.shape[1]


SIDES = {capi.QUAD: [(0, 1), (1, 2), (2, 3), (3, 0)], capi.TRI: [(0, 1), (1, 2), (2, 0)],
         capi.HEX: [(0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 4, 7, 3), (0, 3, 2, 1), (4, 5, 6, 7)]}


def check_numbering(f, b, nodes, seed):
    C, V, dim = nodes.shape
    cv = dpg.vertex_ids(nodes)
    perm = np.random.default_rng(seed).permutation(cv.max() + 1).astype(np.int32)
    cv = perm[cv] * 3 + 5  # relabelled, with gaps: the numbering compacts ids
    gdof, n = dpg.skeleton_numbering(f, b, cv)
    s = dpg.sizes(f, b)
    assert gdof.shape == (C, s.T) and gdof.min() >= 0 and gdof.max() < n
    xi = dpg.skeleton_nodes(f, b)
    Nv = vertex_basis(b.topology, xi)  # (T, V)
    x = np.einsum("tv,cvd->ctd", Nv, nodes)  # physical node of every local trailing dof
    # trailing dofs: (side variable, side) blocks in trial order; traces GLL of
    # degree k per side direction, fluxes degree k - 1
    sides = SIDES[b.topology]
    blocks = [(f.trial_kind[v], v, si) for v in range(f.n_trial) if f.trial_kind[v] in (capi.TRACE, capi.FLUX)
              for si in range(len(sides))]
    ct, cf = (b.h1_order + 1) ** (dim - 1), b.h1_order ** (dim - 1)
    assert sum(ct if k == capi.TRACE else cf for k, _, _ in blocks) == s.T
    key_of = {}
    owner = {}
    for c in range(C):
        off = 0
        for kind, var, si in blocks:
            card = ct if kind == capi.TRACE else cf
            for q in range(card):
                pos = tuple(np.round(x[c, off + q], 9))
                if kind == capi.TRACE:
                    key = (var, pos)
                else:
                    key = (var, tuple(sorted(cv[c, list(sides[si])])), pos)
                g = int(gdof[c, off + q])
                assert key_of.setdefault(g, key) == key, (c, off + q)
                assert owner.setdefault(key, g) == g, (c, off + q)
            off += card
    assert len(key_of) == n  # every global dof is used, one node each
    return n, cv


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_quad_numbering_shares_nodes_exactly(seed):
    f = dpg.form("poisson", 2)
    b = dpg.bucket(capi.QUAD, 4, 5)
    nodes, _ = dpg.mesh_quad(5, 4, 0.15, 40 + seed)
    n, cv = check_numbering(f, b, nodes, seed)
    nv, ne = 6 * 5, 5 * 5 + 6 * 4
    assert n == (nv + ne * (4 - 1)) + ne * 4  # trace: vertices + edge interiors; flux: 4 per edge


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_hex_numbering_shares_nodes_exactly(seed):
    f = dpg.form("poisson", 3)
    b = dpg.bucket(capi.HEX, 3, 4)
    nodes, _ = dpg.mesh_hex(2, 3, 2, 0.1, 50 + seed)
    n, _ = check_numbering(f, b, nodes, seed)
    nv = 3 * 4 * 3
    ne = 2 * 4 * 3 + 3 * 3 * 3 + 3 * 4 * 2
    nf = 3 * 3 * 2 + 2 * 4 * 2 + 2 * 3 * 3
    k = 3
    assert n == (nv + ne * (k - 1) + nf * (k - 1) ** 2) + nf * k * k


def test_triangle_and_stokes_numbering():
    nodes, par = dpg.mesh_quad(3, 4, 0.1, 61)
    tn, _ = split_quads(nodes, par)
    check_numbering(dpg.form("cd", 2), dpg.bucket(capi.TRI, 3, 5), tn, 7)
    check_numbering(dpg.form("stokes", 2), dpg.bucket(capi.QUAD, 3, 4), nodes, 8)


def test_numbering_rejects_bad_input():
    f = dpg.form("poisson", 2)
    b = dpg.bucket(capi.QUAD, 3, 4)
    with pytest.raises(capi.DPGError):
        dpg.skeleton_numbering(f, b, np.array([[0, 1, 1, 2]], dtype=np.int32))
    with pytest.raises(capi.DPGError):
        dpg.skeleton_numbering(f, b, np.array([[0, 1, 2, -3]], dtype=np.int32))


def check_entities(f, b, nodes, seed):
    """dpg_skeleton_entities on a relabelled mesh: a vertex entity holds the
    trace dofs at that vertex (one per trace variable, at its position); on
    every side of every cell the flux dofs and the trace dofs interior to the
    side share one entity, which is an edge (2D) or a face (3D); in 3D the trace
    dofs on a face's edges (not its corners) belong to edge entities."""
    C, V, dim = nodes.shape
    cv = dpg.vertex_ids(nodes)
    perm = np.random.default_rng(seed).permutation(cv.max() + 1).astype(np.int32)
    cv = perm[cv] * 2 + 1
    gdof, n = dpg.skeleton_numbering(f, b, cv)
    ent, ne = dpg.skeleton_entities(f, b, cv, n)
    assert ent.min() >= 0 and ent.max() < ne
    nv = np.unique(cv).size
    xi = dpg.skeleton_nodes(f, b)
    Nv = vertex_basis(b.topology, xi)
    x = np.einsum("tv,cvd->ctd", Nv, nodes)
    sides = SIDES[b.topology]
    blocks = [(f.trial_kind[v], v, si) for v in range(f.n_trial) if f.trial_kind[v] in (capi.TRACE, capi.FLUX)
              for si in range(len(sides))]
    k = b.h1_order
    ct_, cf = (k + 1) ** (dim - 1), k ** (dim - 1)
    ntrace = sum(1 for v in range(f.n_trial) if f.trial_kind[v] == capi.TRACE)
    vpos = {}
    for c in range(C):
        off = 0
        side_ent = {}
        for kind, var, si in blocks:
            card = ct_ if kind == capi.TRACE else cf
            g = gdof[c, off:off + card]
            e = ent[g]
            if kind == capi.FLUX:
                assert np.all(e == e[0]) and e[0] >= nv
                side_ent.setdefault(si, set()).add(int(e[0]))
            else:
                if dim == 2:
                    corner = np.array([0, k])
                    interior = np.arange(1, k)
                    for q in corner:
                        assert e[q] < nv
                        vpos.setdefault(int(e[q]), set()).add(tuple(np.round(x[c, off + q], 9)))
                    side_ent.setdefault(si, set()).update(int(v) for v in e[interior])
                else:
                    a_, b_ = np.meshgrid(np.arange(k + 1), np.arange(k + 1), indexing="xy")
                    a_, b_ = a_.ravel(), b_.ravel()
                    is_corner = ((a_ == 0) | (a_ == k)) & ((b_ == 0) | (b_ == k))
                    on_edge = ((a_ == 0) | (a_ == k) | (b_ == 0) | (b_ == k)) & ~is_corner
                    inner = ~((a_ == 0) | (a_ == k) | (b_ == 0) | (b_ == k))
                    for q in np.flatnonzero(is_corner):
                        assert e[q] < nv
                        vpos.setdefault(int(e[q]), set()).add(tuple(np.round(x[c, off + q], 9)))
                    assert np.all(e[on_edge] >= nv)
                    edge_ids = set(int(v) for v in e[on_edge])
                    side_ent.setdefault(si, set()).update(int(v) for v in e[inner])
                    assert not (edge_ids & set(int(v) for v in e[inner]))  # edges and the face differ
            off += card
        for si, es in side_ent.items():
            assert len(es) == 1, (c, si, es)  # fluxes and interior traces of a side: one entity
    for v, ps in vpos.items():
        assert len(ps) == 1  # a vertex entity sits at one point
    counts = np.bincount(ent[np.isin(ent, list(vpos))], minlength=1)
    assert np.all(counts[list(vpos)] == ntrace)  # one dof per trace variable at a vertex
    return ne


@pytest.mark.parametrize("seed", [1, 2])
def test_entities_quads_and_stokes(seed):
    nodes, _ = dpg.mesh_quad(4, 3, 0.1, 80 + seed)
    check_entities(dpg.form("poisson", 2), dpg.bucket(capi.QUAD, 4, 5), nodes, seed)
    check_entities(dpg.form("stokes", 2), dpg.bucket(capi.QUAD, 3, 4), nodes, seed + 10)


def test_entities_hexes():
    nodes, _ = dpg.mesh_hex(2, 2, 3, 0.1, 90)
    ne = check_entities(dpg.form("poisson", 3), dpg.bucket(capi.HEX, 3, 4), nodes, 5)
    nv, ned = 3 * 3 * 4, 2 * 3 * 4 + 3 * 2 * 4 + 3 * 3 * 3
    nf = 2 * 2 * 4 + 2 * 3 * 3 + 3 * 2 * 3
    assert ne == nv + ned + nf
