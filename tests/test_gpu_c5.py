"""The differentiation workload on its own mesh: 2D convection-diffusion-
reaction ultraweak DPG on 64x64 seeded-jittered quads (4,096 cells), p=2,
test order 4, FP64, per-cell coefficients eps log-uniform on [1e-3, 1],
beta = unit vector at a seeded angle times U[0.5, 2], reaction U[0, 1].

Every cell: K, rhs and the Schur complement against the oracle (80-bit floor
rule), the energy error of seeded trial coefficients against the oracle's
(TSolution::rankLocalEnergyError, Solution.cpp:2653-2685), and the adjoint
gradient of the summed indicator with respect to all 4 x 4,096 coefficients
against central differences with step 1e-6 (and a Richardson-extrapolated
reference), once from a fresh factorisation and once from the one the
forward call kept."""
import numpy as np
import pytest

from camellia2_amd import dpg
from oracle import oracle

from helpers import cdr2d, rel_fro

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.fixture(scope="module")
def c5():
    f, b, nodes, par, coefs, load = cdr2d(64, 64)
    s = dpg.sizes(f, b)
    C = nodes.shape[0]
    u = dpg.rng_normal(0xC0FFEE + 55, C * s.M).reshape(C, s.M)
    A = dpg.LocalAssembler(f, [b])
    yield f, b, nodes, par, coefs, load, u, A
    A.close()


def test_c5_forward_every_cell_against_the_oracle(c5):
    f, b, nodes, par, coefs, load, u, A = c5
    out, nbad = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), dev(coefs), load, want_schur=True)
    torch.cuda.synchronize()
    assert nbad == 0
    o, x = oracle.with_extended(oracle.form_like(f), b, nodes, par, coefs, load, schur=True)
    for gk, ok in (("K", "K"), ("rhs", "rhs"), ("schur", "S"), ("schur_rhs", "g")):
        g = out[gk].cpu().numpy()
        floor = rel_fro(o[ok], x[ok])
        err_x, err_o = rel_fro(g, x[ok]), rel_fro(g, o[ok])
        print(f"{gk:10s} gpu-vs-oracle {err_o:.2e}  gpu-vs-80bit {err_x:.2e}  oracle-vs-80bit {floor:.2e}")
        assert err_x <= max(1e-12, 4.0 * floor), gk


def test_c5_energy_error_against_the_oracle(c5):
    f, b, nodes, par, coefs, load, u, A = c5
    dn, dp, dc, du = dev(nodes), dev(par), dev(coefs), dev(u)
    eta2 = A.energy_error(0, dn, dp, du, dc, load).cpu().numpy()
    out, nbad = A.local_stiffness_and_rhs(0, dn, dp, dc, load, want_factor=True)
    assert nbad == 0
    eta2f = A.energy_error_factored(0, out["factor"], du).cpu().numpy()
    ff = oracle.form_like(f)
    ref = oracle.energy_error(ff, b, nodes, par, u, coefs, load)
    x = oracle.eta2_extended(ff, b, nodes, par, u, coefs, load)
    rel = lambda a, r: float(np.max(np.abs(a - r) / np.abs(r)))  # noqa: E731
    floor = rel(ref, x)
    print(f"eta2 vs oracle: re-assembled {rel(eta2, ref):.2e}, from the kept factor {rel(eta2f, ref):.2e}; "
          f"vs 80-bit: {rel(eta2, x):.2e}, {rel(eta2f, x):.2e}; oracle vs 80-bit {floor:.2e}")
    # the FP64 floor rule of the other outputs: G is as ill conditioned here as
    # for K (h = 1/64, eps down to 1e-3), so the oracle's own eta2 is ~4e-10 from
    # the 80-bit evaluation
    assert rel(eta2, x) <= max(1e-11, 4.0 * floor)
    assert rel(eta2f, x) <= max(1e-11, 4.0 * floor)


def test_c5_adjoint_against_central_differences(c5):
    f, b, nodes, par, coefs, load, u, A = c5
    dn, dp, dc, du = dev(nodes), dev(par), dev(coefs), dev(u)
    eta2, grad = A.energy_error_and_gradient(0, dn, dp, du, dc, load)
    out, nbad = A.local_stiffness_and_rhs(0, dn, dp, dc, load, want_factor=True)
    assert nbad == 0
    grad_kept = A.adjoint(0, dn, dp, dc, load, u=du, factor=out["factor"], factor_scale=out["factor_scale"])
    grad, grad_kept = grad.cpu().numpy(), grad_kept.cpu().numpy()
    # the two paths factor G in different kernel sequences (the assembly's
    # autotuned tiles against the adjoint's own factorisation), so they agree to
    # the conditioning of G at h = 1/64, eps down to 1e-3 (the same ~1e-10 that
    # separates the oracle from the 80-bit evaluation), not to 1e-12
    assert np.linalg.norm(grad_kept - grad) <= 1e-10 * np.linalg.norm(grad)

    def eta(c):
        return A.energy_error(0, dn, dp, du, dev(c), load).cpu().numpy()

    check_against_differences(eta, coefs, grad, "adjoint")


def check_against_differences(fn, coefs, grad, what):
    """The gradient of sum(fn) against central differences with step 1e-6 and
    against the Richardson extrapolation of two central differences (steps
    5e-3 and 1e-2, relative to eps for slot 0, absolute for the others), whose
    own uncertainty is its distance from the extrapolation with half the steps.
    At h = 1/64 and eps down to 1e-3 fn carries ~1e-10 relative rounding (the
    oracle's own distance from the 80-bit evaluation), which the step-1e-6
    quotient turns into ~1e-6..1e-5 of noise: the adjoint must lie within 1e-6
    of the extrapolated derivative (or within twice that reference's own
    uncertainty where the truncation of the extrapolation is larger) and no
    farther from the step-1e-6 quotient than that quotient's own noise plus
    1e-6."""
    rows = []
    for k in range(coefs.shape[1]):
        scale = coefs[:, k] if k == 0 else np.ones(coefs.shape[0])

        def central(r):
            hh = r * scale
            cp, cm = coefs.copy(), coefs.copy()
            cp[:, k] += hh
            cm[:, k] -= hh
            return (fn(cp) - fn(cm)) / (2 * hh)

        fd = central(1e-6 / scale)
        c1, c2, c4 = central(2.5e-3), central(5e-3), central(1e-2)
        rich, rich_half = (4.0 * c2 - c4) / 3.0, (4.0 * c1 - c2) / 3.0
        nr = np.linalg.norm(rich)
        e_fd = np.linalg.norm(grad[:, k] - fd) / np.linalg.norm(fd)
        e_rich = np.linalg.norm(grad[:, k] - rich) / nr
        e_ref = np.linalg.norm(rich - rich_half) / nr
        e_fd_rich = np.linalg.norm(fd - rich) / nr
        print(f"coef {k}: {what} vs central differences (h=1e-6) {e_fd:.2e}, vs Richardson {e_rich:.2e}; "
              f"the h=1e-6 quotient is {e_fd_rich:.2e} from Richardson, Richardson {e_ref:.2e} from its half step")
        rows.append((k, e_fd, e_rich, e_ref, e_fd_rich))
    for k, e_fd, e_rich, e_ref, e_fd_rich in rows:
        assert e_rich <= max(1e-6, 2.0 * e_ref), (k, e_rich, e_ref)
        assert e_fd <= e_fd_rich + 1e-6, (k, e_fd, e_fd_rich)


def test_c5_adjoint_of_K_and_rhs_against_central_differences(c5):
    f, b, nodes, par, coefs, load, u, A = c5
    s = A.info[0]
    C = nodes.shape[0]
    dn, dp, dc = dev(nodes), dev(par), dev(coefs)
    Kb = dev(dpg.rng_normal(0xC0FFEE + 56, C * s.M * s.M).reshape(C, s.M, s.M))
    rb = dev(dpg.rng_normal(0xC0FFEE + 57, C * s.M).reshape(C, s.M))
    out, nbad = A.local_stiffness_and_rhs(0, dn, dp, dc, load, want_factor=True)
    assert nbad == 0
    grad = A.adjoint(0, dn, dp, dc, load, Kbar=Kb, rhsbar=rb, factor=out["factor"],
                     factor_scale=out["factor_scale"]).cpu().numpy()

    def loss(c):
        o, nb = A.local_stiffness_and_rhs(0, dn, dp, dev(c), load)
        return ((o["K"] * Kb).sum(dim=(1, 2)) + (o["rhs"] * rb).sum(dim=1)).cpu().numpy()

    check_against_differences(loss, coefs, grad, "adjoint of <Kbar,K> + <rbar,rhs>")
