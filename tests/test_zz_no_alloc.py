"""dpg_assemble_noalloc, the entry point that allocates no device memory (its
workspace is the library's static device scratch), against dpg_assemble on the
same plan: the same bits for every output and the same INFO, on two calls in a
row with the second call smaller, for a 2D form with per-cell coefficients, the
Stokes form and hexes at the 32^3-workload orders (several scratch-sized
batches per call). Also: the plan's own workspace stays unallocated, and an
element type too large for the scratch is refused with a message."""
import ctypes as ct

import numpy as np
import pytest

from camellia2_amd import capi, dpg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KEYS = ("K", "rhs", "schur", "schur_rhs", "info")


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def noalloc(A, nodes, par, coefs, load):
    """dpg_assemble_noalloc through the C-ABI into fresh outputs (NaN-filled, so
    that an output left unwritten cannot pass)"""
    C = nodes.shape[0]
    i = A.info[0]
    o = {"K": torch.full((C, i.M, i.M), float("nan"), dtype=torch.float64, device="cuda:0"),
         "rhs": torch.full((C, i.M), float("nan"), dtype=torch.float64, device="cuda:0"),
         "schur": torch.full((C, i.T, i.T), float("nan"), dtype=torch.float64, device="cuda:0"),
         "schur_rhs": torch.full((C, i.T), float("nan"), dtype=torch.float64, device="cuda:0"),
         "info": torch.full((C,), -7, dtype=torch.int32, device="cuda:0")}
    _, outs = A._outputs(0, C, nodes.device, o, False, True, True)
    b = A._batch(nodes, par, coefs, load)
    s = torch.cuda.current_stream().cuda_stream
    rc = capi.lib().dpg_assemble_noalloc(A._plan, 0, ct.byref(b), ct.byref(outs), ct.c_void_p(s))
    torch.cuda.synchronize()
    assert rc >= 0, capi.lib().dpg_last_error().decode()
    return o, rc


def check_pair(A, nodes, par, coefs, load, calls):
    """each (lo, hi) of `calls` in turn: noalloc on cells [lo, hi) equals dpg_assemble on them"""
    for lo, hi in calls:
        n_, p_ = dev(nodes[lo:hi]), dev(par[lo:hi])
        c_ = dev(coefs[lo:hi]) if coefs is not None else None
        got, bad = noalloc(A, n_, p_, c_, load)
        want, bad_w = A.local_stiffness_and_rhs(0, n_, p_, c_, load, want_schur=True, want_info=True)
        torch.cuda.synchronize()
        assert bad == bad_w == 0
        for k in KEYS:
            assert torch.equal(got[k], want[k]), (lo, hi, k)


def test_noalloc_equals_assemble_cdr_2d_with_coefficients():
    f = dpg.form("cdr", 2)
    b = dpg.bucket(capi.QUAD, 3, 4)
    nodes, par = dpg.mesh_quad(16, 16, 0.1, 0xC0FFEE + 5)
    C = nodes.shape[0]
    coefs = np.stack([10.0 ** dpg.rng_uniform(71, -3.0, 0.0, C), dpg.rng_uniform(72, 0.5, 2.0, C),
                      dpg.rng_uniform(73, -1.0, 1.0, C), dpg.rng_uniform(74, 0.0, 1.0, C)], axis=1)
    load = dpg.load_sine_modes(2, 1, 8, 6, 0xC0FFEE + 101)
    A = dpg.LocalAssembler(f, [b])
    # the first call on a fresh plan: its own workspace is never allocated
    n_, p_, c_ = dev(nodes[:100]), dev(par[:100]), dev(coefs[:100])
    noalloc(A, n_, p_, c_, load)
    assert A.workspace_bytes() == 0
    check_pair(A, nodes, par, coefs, load, [(0, C), (37, 60)])
    A.close()


def test_noalloc_equals_assemble_stokes():
    f = dpg.form("stokes", 2)
    b = dpg.bucket(capi.QUAD, 4, 5)
    nodes, par = dpg.mesh_quad(12, 12, 0.1, 0xC0FFEE + 2)
    load = dpg.load_sine_modes(2, 2, 8, 6, 0xC0FFEE + 102)
    A = dpg.LocalAssembler(f, [b])
    check_pair(A, nodes, par, None, load, [(0, 144), (5, 18)])
    A.close()


def test_noalloc_equals_assemble_hex_several_batches():
    """32^3-workload orders (N = 666, M = 502): a cell needs ~27 MB of workspace,
    so the 256 MiB scratch takes 9 cells per batch and 20 cells run as 3 batches"""
    f = dpg.form("poisson", 3)
    b = dpg.bucket(capi.HEX, 4, 5)
    nodes, par = dpg.mesh_hex(4, 3, 2, 0.05, 0xC0FFEE + 4)
    load = dpg.load_sine_modes(3, 1, 16, 6, 0xC0FFEE + 104)
    A = dpg.LocalAssembler(f, [b])
    check_pair(A, nodes, par, None, load, [(0, 20), (3, 10)])
    A.close()


def test_noalloc_refuses_an_element_type_larger_than_the_scratch():
    """hexes with H1 order 7 and test order 9: N = 3430, M = 2050, one cell's
    augmented matrix alone is ~240 MB, so with the Gram operand it exceeds the
    256 MiB scratch; the call fails before any launch and says why"""
    f = dpg.form("poisson", 3)
    b = dpg.bucket(capi.HEX, 7, 9)
    nodes, par = dpg.mesh_hex(1, 1, 1, 0.0, 1)
    A = dpg.LocalAssembler(f, [b])
    i = A.info[0]
    o = {"K": torch.empty((1, i.M, i.M), dtype=torch.float64, device="cuda:0"),
         "rhs": torch.empty((1, i.M), dtype=torch.float64, device="cuda:0")}
    _, outs = A._outputs(0, 1, torch.device("cuda:0"), o, False, False, False)
    bt = A._batch(dev(nodes), dev(par), None, dpg.load_sine_modes(3, 1, 1, 1, 5))
    rc = capi.lib().dpg_assemble_noalloc(A._plan, 0, ct.byref(bt), ct.byref(outs),
                                         ct.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc < 0
    msg = capi.lib().dpg_last_error().decode()
    assert "static scratch" in msg and "dpg_assemble" in msg
    A.close()
