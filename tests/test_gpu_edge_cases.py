"""Batch-shape edge cases on the GPU: an empty batch, single cells, ragged
batch splits and a call with more cells than one grid dimension holds. A cell's
outputs depend on that cell's inputs only, so they must be the same bits
whatever batch the cell is assembled in."""
import numpy as np
import pytest

from camellia2_amd import capi, dpg
from oracle import oracle

from helpers import rel_fro

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KEYS = ("K", "rhs", "schur", "schur_rhs")


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def cdr_case(nx, ny, seed):
    """per-cell coefficients (eps, beta_x, beta_y, alpha), so every cell differs"""
    f = dpg.form("cdr", 2)
    b = dpg.bucket(capi.QUAD, 3, 4)
    nodes, par = dpg.mesh_quad(nx, ny, 0.1, seed)
    C = nodes.shape[0]
    coefs = np.stack([10.0 ** dpg.rng_uniform(seed + 1, -3.0, 0.0, C), dpg.rng_uniform(seed + 2, 0.5, 2.0, C),
                      dpg.rng_uniform(seed + 3, -1.0, 1.0, C), dpg.rng_uniform(seed + 4, 0.0, 1.0, C)], axis=1)
    load = dpg.load_sine_modes(2, 1, 8, 6, seed + 100)
    return f, b, nodes, par, coefs, load


def test_empty_batch_does_nothing_and_reports_zero():
    f, b, nodes, par, coefs, load = cdr_case(4, 4, 11)
    A = dpg.LocalAssembler(f, [b])
    out, nbad = A.local_stiffness_and_rhs(0, dev(nodes[:0]), dev(par[:0]), dev(coefs[:0]), load, want_schur=True)
    torch.cuda.synchronize()
    assert nbad == 0 and out["K"].shape[0] == 0 and out["schur"].shape[0] == 0
    u0 = dev(np.zeros((0, A.info[0].M)))
    eta2 = A.energy_error(0, dev(nodes[:0]), dev(par[:0]), u0, dev(coefs[:0]), load)
    eta2g, grad = A.energy_error_and_gradient(0, dev(nodes[:0]), dev(par[:0]), u0, dev(coefs[:0]), load)
    assert eta2.shape == (0,) and eta2g.shape == (0,) and grad.shape == (0, 4)
    A.close()


def test_single_cells_and_ragged_batches_give_the_same_bits():
    f, b, nodes, par, coefs, load = cdr_case(9, 7, 21)  # 63 cells
    A = dpg.LocalAssembler(f, [b])
    full, nbad = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), dev(coefs), load, want_schur=True)
    assert nbad == 0
    for c in (0, 31, 62):
        one, nbad = A.local_stiffness_and_rhs(0, dev(nodes[c:c + 1]), dev(par[c:c + 1]), dev(coefs[c:c + 1]), load,
                                              want_schur=True)
        assert nbad == 0
        for k in KEYS:
            assert torch.equal(one[k][0], full[k][c]), (c, k)
    A.close()
    # the same cells through batches of 5 (12 full batches and one of 3)
    Ar = dpg.LocalAssembler(f, [b], max_batch_cells=5)
    rag, nbad = Ar.local_stiffness_and_rhs(0, dev(nodes), dev(par), dev(coefs), load, want_schur=True)
    assert nbad == 0
    for k in KEYS:
        assert torch.equal(rag[k], full[k]), k
    Ar.close()
    # and one of them against the oracle (80-bit floor rule)
    c = 31
    o, x = oracle.with_extended(oracle.form_like(f), b, nodes[c:c + 1], par[c:c + 1], coefs[c:c + 1], load)
    err = rel_fro(full["K"][c:c + 1].cpu().numpy(), x["K"])
    assert err <= max(1e-12, 4.0 * rel_fro(o["K"], x["K"]))


def test_more_cells_than_one_grid_dimension():
    """70,000 cells in one call: the library splits the call into batches of at
    most 65,535 cells (a grid dimension); every cell equals its value from a
    call split into 10,007-cell batches, and sampled cells match the oracle."""
    f, b, nodes, par, coefs, load = cdr_case(280, 250, 31)
    C = nodes.shape[0]
    assert C == 70000
    big = dpg.LocalAssembler(f, [b], workspace_bytes=96.0 * 2 ** 30)
    d = (dev(nodes), dev(par), dev(coefs))
    out, nbad = big.local_stiffness_and_rhs(0, *d, load, want_schur=True)
    assert nbad == 0
    big.close()
    small = dpg.LocalAssembler(f, [b], max_batch_cells=10007)
    ref, nbad = small.local_stiffness_and_rhs(0, *d, load, want_schur=True)
    assert nbad == 0
    small.close()
    for k in KEYS:
        assert torch.equal(out[k], ref[k]), k
    sel = np.array([0, 65534, 65535, 69999])  # both sides of the split
    o, x = oracle.with_extended(oracle.form_like(f), b, nodes[sel], par[sel], coefs[sel], load)
    err = rel_fro(out["K"][torch.from_numpy(sel).to("cuda:0")].cpu().numpy(), x["K"])
    assert err <= max(1e-12, 4.0 * rel_fro(o["K"], x["K"]))
