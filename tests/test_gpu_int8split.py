"""Precision DPG_PREC_INT8SPLIT: the Gram product on int8 MFMA from 8 signed
7-bit digits of the equilibrated operand (exact int32 digit products, levels
summed in FP64), everything else FP64. On patches two bounds: the
mixed-precision target (K, rhs and the condensed system within 1e-10 relative
of the FP64 oracle) and a floor-relative one -- the distance from the 80-bit
evaluation at most max(1e-11, 8 * floor), floor the reference algorithm's own
FP64 distance from it on the same cells (the plain FP64 form's bar is
max(1e-12, 4 * floor)). On the fine mesh see test_int8split_full_mesh_cells."""
import numpy as np
import pytest

from camellia2_amd import capi, dpg
from oracle import oracle

from helpers import cdr2d, poisson2d, poisson3d, rel_fro, split_quads, stokes2d

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL_MIXED = 1e-10  # mixed-precision target, relative to the FP64 oracle
I8 = capi.PREC_INT8SPLIT


def floor_bound(floor):
    return max(1e-11, 8.0 * floor)


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def assemble(case, precision):
    f, b, nodes, par, coefs, load = case
    A = dpg.LocalAssembler(f, [b], precision=precision)
    out, nbad = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), dev(coefs), load, want_schur=True)
    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    A.close()
    return res, nbad


def tri_case():
    f, b, nodes, par, coefs, load = cdr2d(4, 3, k=3, kt=5)
    tn, tp = split_quads(nodes, par)
    tc = np.repeat(coefs, 2, axis=0)
    return f, dpg.bucket(capi.TRI, 3, 5), tn, tp, tc, load


CASES = {
    "poisson2d": lambda: poisson2d(4, 4),
    "cdr2d": lambda: cdr2d(4, 4),
    "stokes2d": lambda: stokes2d(3, 3),
    "triangles": tri_case,
    "poisson3d": lambda: poisson3d(2, 2, 1),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_int8split_meets_the_fp64_bar(name):
    case = CASES[name]()
    f, b, nodes, par, coefs, load = case
    g, nbad = assemble(case, I8)
    g64, _ = assemble(case, capi.PREC_FP64)
    assert nbad == 0 and np.all(g["info"] == 0)
    o, x = oracle.with_extended(oracle.form_like(f), b, nodes, par, coefs, load, schur=True)
    for gk, ok in (("K", "K"), ("rhs", "rhs"), ("schur", "S"), ("schur_rhs", "g")):
        floor = rel_fro(o[ok], x[ok])
        err_x = rel_fro(g[gk], x[ok])
        err_o = rel_fro(g[gk], o[ok])
        e64 = rel_fro(g64[gk], x[ok])
        print(f"{name} {gk:10s} int8split-vs-oracle {err_o:.2e}  int8split-vs-80bit {err_x:.2e}  "
              f"fp64-vs-80bit {e64:.2e}  oracle-vs-80bit {floor:.2e}")
        assert err_o <= TOL_MIXED, gk
        assert err_x <= floor_bound(floor), gk
    assert np.array_equal(g["K"], np.transpose(g["K"], (0, 2, 1)))
    assert not np.array_equal(g["K"], g64["K"])  # the int8 path really ran


TOL_FINE_MESH = 5e-9


def test_int8split_full_mesh_cells():
    """Seeded cells of the 32^3 hex mesh (h = 1/32, the worst-conditioned G: the
    reference algorithm's own FP64 K is ~6e-10 from the 80-bit evaluation). The
    split's truncation (dropped digit levels, ~2^-53 of the column norms per
    term) is amplified by G's conditioning like any perturbation of G: K, rhs
    and Schur come out ~1e-9 from the 80-bit result here (the FP64 form:
    ~6e-10 for K, ~4e-11 for rhs), bounded by TOL_FINE_MESH."""
    f, b, nodes, par, coefs, load = poisson3d(32, 32, 32)
    sample = np.sort(np.random.default_rng(4).choice(nodes.shape[0], 4, replace=False))
    case = (f, b, nodes[sample], par[sample], None, load)
    g, nbad = assemble(case, I8)
    assert nbad == 0
    o, x = oracle.with_extended(oracle.form_like(f), b, nodes[sample], par[sample], None, load, schur=True)
    for gk, ok in (("K", "K"), ("rhs", "rhs"), ("schur", "S")):
        floor = rel_fro(o[ok], x[ok])
        err_x = rel_fro(g[gk], x[ok])
        print(f"32^3 cells {gk:6s} int8split-vs-80bit {err_x:.2e}  oracle-vs-80bit {floor:.2e}")
        assert err_x <= TOL_FINE_MESH, gk


def test_int8split_energy_error_and_batches():
    """The energy error through the int8 Gram, and batching invariance."""
    f, b, nodes, par, coefs, load = cdr2d(5, 4)
    s = dpg.sizes(f, b)
    u = np.asarray(dpg.rng_normal(91, nodes.shape[0] * s.M)).reshape(nodes.shape[0], s.M)
    A = dpg.LocalAssembler(f, [b], precision=I8)
    eta2 = A.energy_error(0, dev(nodes), dev(par), dev(u), dev(coefs), load).cpu().numpy()
    A.close()
    ref = oracle.energy_error(oracle.form_like(f), b, nodes, par, u, coefs, load)
    err = np.max(np.abs(eta2 - ref) / np.abs(ref))
    print(f"energy error vs oracle: {err:.2e}")
    assert err < 1e-11
    g1, _ = assemble((f, b, nodes, par, coefs, load), I8)
    A = dpg.LocalAssembler(f, [b], precision=I8, max_batch_cells=3)
    out, nbad = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), dev(coefs), load, want_schur=True)
    K3 = out["K"].cpu().numpy()
    A.close()
    assert nbad == 0 and np.array_equal(K3, g1["K"])


def test_int8split_rejects_gram_output():
    f, b, nodes, par, coefs, load = poisson2d(2, 2)
    A = dpg.LocalAssembler(f, [b], precision=I8)
    with pytest.raises(capi.DPGError):
        A.local_stiffness_and_rhs(0, dev(nodes), dev(par), None, load, want_G=True)
    A.close()
