"""Input preparation on the device: the seeded jittered meshes generated on
the device (bit for bit the host generators), side parities of conforming
meshes from their vertex ids (the generators' parities, the reference's rule),
and element-type bucketing with the gather of the cells' inputs."""
import numpy as np
import pytest

from camellia2_amd import capi, dpg

from helpers import cd_hp, split_quads

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("kind", ["quad", "hex", "tri"])
def test_side_parities_match_the_generators(kind):
    if kind == "quad":
        nodes, par = dpg.mesh_quad(7, 5, 0.1, 131)
        topo = capi.QUAD
    elif kind == "hex":
        nodes, par = dpg.mesh_hex(3, 2, 4, 0.05, 132)
        topo = capi.HEX
    else:
        n4, p4 = dpg.mesh_quad(4, 3, 0.1, 133)
        nodes, par = split_quads(n4, p4)
        topo = capi.TRI
    cv = dpg.vertex_ids(nodes)
    got = dpg.side_parities(topo, dev(cv)).cpu().numpy()
    assert np.array_equal(got, par)
    # vertex ids do not enter the rule, only which cells share a side
    perm = np.random.default_rng(5).permutation(cv.max() + 1).astype(np.int32)
    assert np.array_equal(dpg.side_parities(topo, dev(perm[cv])).cpu().numpy(), par)


def test_side_shared_by_three_cells_is_rejected():
    nodes, _ = dpg.mesh_quad(3, 1, 0.0, 1)
    cv = dpg.vertex_ids(nodes)
    bad = np.concatenate([cv, cv[1:2]])  # a copy of the middle cell: its sides now have three cells
    with pytest.raises(capi.DPGError):
        dpg.side_parities(capi.QUAD, dev(bad))


def test_bucketing_and_gather_match_numpy():
    rng = np.random.default_rng(7)
    C = 5000
    keys = rng.integers(0, 9, C).astype(np.int32)
    keys[keys == 4] = 11  # a gap in the key values
    perm, offsets, bkeys = dpg.sort_cells(dev(keys))
    want = np.argsort(keys, kind="stable")
    assert np.array_equal(perm.cpu().numpy(), want)
    uk, first = np.unique(keys[want], return_index=True)
    assert np.array_equal(bkeys.cpu().numpy(), uk)
    assert np.array_equal(offsets.cpu().numpy(), np.append(first, C))
    nodes = rng.normal(size=(C, 4, 2))
    par = rng.choice([-1, 1], size=(C, 4)).astype(np.int8)
    coefs = rng.normal(size=(C, 3))
    n2, p2, c2 = dpg.gather_cells(perm, dev(nodes), dev(par), dev(coefs))
    assert np.array_equal(n2.cpu().numpy(), nodes[want])
    assert np.array_equal(p2.cpu().numpy(), par[want])
    assert np.array_equal(c2.cpu().numpy(), coefs[want])


def test_bucketing_the_hp_mesh_by_element_type():
    """The graded hp mesh's cells grouped by (topology, p, dp) on the device give
    the same element types and cell counts as the host grouping."""
    f, cells, load = cd_hp(4, 4, tri_cols=1, tri_rows=0)
    q = cells[capi.QUAD]
    key = (q["p"] * 8 + q["dp"]).astype(np.int32)
    perm, offsets, bkeys = dpg.sort_cells(dev(key))
    counts = np.diff(offsets.cpu().numpy())
    uk, cnt = np.unique(key, return_counts=True)
    assert np.array_equal(bkeys.cpu().numpy(), uk) and np.array_equal(counts, cnt)
    n2, p2, c2 = dpg.gather_cells(perm, dev(q["nodes"]), dev(q["parities"]), dev(q["coefs"]))
    o = offsets.cpu().numpy()
    for b, k in enumerate(uk):
        sel = np.flatnonzero(key == k)
        assert np.array_equal(n2[o[b]:o[b + 1]].cpu().numpy(), q["nodes"][sel])
        assert np.array_equal(c2[o[b]:o[b + 1]].cpu().numpy(), q["coefs"][sel])


@pytest.mark.parametrize("shape", [(1, 1), (2, 1), (7, 5), (16, 16), (64, 64), (128, 128), (301, 17)])
def test_device_quad_mesh_equals_the_host_generator(shape):
    nx, ny = shape
    for seed, jitter in ((0xC0FFEE + 1, 0.10), (12345, 0.49), (7, 0.0)):
        hn, hp = dpg.mesh_quad(nx, ny, jitter, seed)
        dn, dp = dpg.mesh_quad_device(nx, ny, jitter, seed)
        torch.cuda.synchronize()
        assert np.array_equal(dn.cpu().numpy(), hn), (shape, seed)
        assert np.array_equal(dp.cpu().numpy(), hp), (shape, seed)


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 4), (5, 4, 3), (32, 32, 32), (40, 7, 9)])
def test_device_hex_mesh_equals_the_host_generator(shape):
    hn, hp = dpg.mesh_hex(*shape, 0.05, 0xC0FFEE + 4)
    dn, dp = dpg.mesh_hex_device(*shape, 0.05, 0xC0FFEE + 4)
    torch.cuda.synchronize()
    assert np.array_equal(dn.cpu().numpy(), hn), shape
    assert np.array_equal(dp.cpu().numpy(), hp), shape


def test_device_mesh_rejects_bad_arguments():
    with pytest.raises(capi.DPGError):
        dpg.mesh_quad_device(0, 4, 0.1, 1)
    with pytest.raises(capi.DPGError):
        dpg.mesh_hex_device(2, 2, 2, 0.7, 1)
