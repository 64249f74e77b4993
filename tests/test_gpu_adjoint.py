"""Adjoint of the local assembly with respect to the per-cell coefficients
(dpg_adjoint with the cotangents Kbar, rhsbar and the energy-error term),
against central differences of the GPU forward, with and without the
factorisation kept by dpg_assemble; the energy error from that kept
factorisation (dpg_energy_error_factored) against dpg_energy_error and the
oracle; and the torch.autograd wrapper over (K, rhs)."""
import numpy as np
import pytest

from camellia2_amd import dpg
from oracle import oracle

from helpers import cdr2d

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def case(n=6, seed=0xC0FFEE + 55):
    f, b, nodes, par, coefs, load = cdr2d(n, n, k=3, kt=4)  # C5 orders: p=2, test 4
    s = dpg.sizes(f, b)
    C = nodes.shape[0]
    u = dpg.rng_normal(seed, C * s.M).reshape(C, s.M)
    Kb = dpg.rng_normal(seed + 1, C * s.M * s.M).reshape(C, s.M, s.M)
    rb = dpg.rng_normal(seed + 2, C * s.M).reshape(C, s.M)
    return f, b, nodes, par, coefs, load, u, Kb, rb


def kr_loss(A, dn, dp, coefs, load, Kb, rb):
    """per cell <Kbar, K> + <rhsbar, rhs> of the GPU forward"""
    out, nbad = A.local_stiffness_and_rhs(0, dn, dp, dev(coefs), load)
    assert nbad == 0
    return ((out["K"] * Kb).sum(dim=(1, 2)) + (out["rhs"] * rb).sum(dim=1)).cpu().numpy()


def richardson(fun, coefs, k):
    """d fun / d coefs[:, k] per cell: Richardson extrapolation of central
    differences with steps 1e-3 and 5e-4 (relative for eps, which enters as
    sqrt and 1/eps; absolute for beta and the reaction)"""
    def central(rel):
        hh = rel * (coefs[:, k] if k == 0 else np.ones(coefs.shape[0]))
        cp, cm = coefs.copy(), coefs.copy()
        cp[:, k] += hh
        cm[:, k] -= hh
        return (fun(cp) - fun(cm)) / (2 * hh)

    return (4.0 * central(5e-4) - central(1e-3)) / 3.0


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# The differences of <Kbar, K> carry the rounding of K (~1e-12 of |Kbar||K|,
# a sum over M^2 entries of random sign) divided by the 1e-3 step: up to ~1e-8
# of the derivative for a coefficient K depends on weakly (the reaction). A
# wrong term or sign in the adjoint gives O(1) differences.
TOL_KR = 1e-7


def test_adjoint_of_K_and_rhs_matches_finite_differences():
    f, b, nodes, par, coefs, load, u, Kb, rb = case()
    A = dpg.LocalAssembler(f, [b])
    dn, dp, dKb, drb = dev(nodes), dev(par), dev(Kb), dev(rb)
    grad = A.adjoint(0, dn, dp, dev(coefs), load, Kbar=dKb, rhsbar=drb).cpu().numpy()
    for k in range(coefs.shape[1]):
        ref = richardson(lambda c: kr_loss(A, dn, dp, c, load, dKb, drb), coefs, k)
        e = rel(grad[:, k], ref)
        print(f"coef {k}: adjoint of <Kbar,K> + <rbar,rhs> vs Richardson differences {e:.2e}")
        assert e < TOL_KR
    # each cotangent on its own
    for kb, rb_, name in ((dKb, None, "Kbar"), (None, drb, "rhsbar")):
        g1 = A.adjoint(0, dn, dp, dev(coefs), load, Kbar=kb, rhsbar=rb_).cpu().numpy()
        for k in range(coefs.shape[1]):
            ref = richardson(lambda c: kr_loss(A, dn, dp, c, load, kb if kb is not None else torch.zeros_like(dKb),
                                               rb_ if rb_ is not None else torch.zeros_like(drb)), coefs, k)
            e = rel(g1[:, k], ref)
            print(f"coef {k}: {name} alone {e:.2e}")
            assert e < TOL_KR
    A.close()


def test_adjoint_from_the_kept_factorisation():
    """The factorisation dpg_assemble keeps gives the same gradient bits as the
    adjoint that factors itself (no u: both start from the same columns), and
    the sum of the three parts is the gradient of their sum."""
    f, b, nodes, par, coefs, load, u, Kb, rb = case(5)
    A = dpg.LocalAssembler(f, [b])
    dn, dp, dc, du, dKb, drb = dev(nodes), dev(par), dev(coefs), dev(u), dev(Kb), dev(rb)
    out, nbad = A.local_stiffness_and_rhs(0, dn, dp, dc, load, want_factor=True)
    assert nbad == 0
    fac, fs = out["factor"], out["factor_scale"]
    g_call = A.adjoint(0, dn, dp, dc, load, Kbar=dKb, rhsbar=drb)
    g_kept = A.adjoint(0, dn, dp, dc, load, Kbar=dKb, rhsbar=drb, factor=fac, factor_scale=fs)
    assert torch.equal(g_call, g_kept)
    # with u: z comes from the factorisation's residual row in one, from l - B u
    # in the kept factor's coordinates in the other (same value up to rounding)
    e_call, ga = A.adjoint(0, dn, dp, dc, load, u=du, Kbar=dKb, rhsbar=drb, want_eta2=True)
    e_kept, gb = A.adjoint(0, dn, dp, dc, load, u=du, Kbar=dKb, rhsbar=drb, factor=fac, factor_scale=fs,
                           want_eta2=True)
    assert rel(e_kept.cpu().numpy(), e_call.cpu().numpy()) < 1e-13
    assert rel(gb.cpu().numpy(), ga.cpu().numpy()) < 1e-12
    parts = (A.adjoint(0, dn, dp, dc, load, Kbar=dKb) + A.adjoint(0, dn, dp, dc, load, rhsbar=drb) +
             A.adjoint(0, dn, dp, dc, load, u=du))
    assert rel(parts.cpu().numpy(), ga.cpu().numpy()) < 1e-12
    A.close()


def test_energy_error_from_the_kept_factorisation():
    f, b, nodes, par, coefs, load, u, _, _ = case(8)
    A = dpg.LocalAssembler(f, [b])
    dn, dp, dc, du = dev(nodes), dev(par), dev(coefs), dev(u)
    out, nbad = A.local_stiffness_and_rhs(0, dn, dp, dc, load, want_factor=True)
    assert nbad == 0
    eta_f = A.energy_error_factored(0, out["factor"], du).cpu().numpy()
    eta_e = A.energy_error(0, dn, dp, du, dc, load).cpu().numpy()
    ref = oracle.energy_error(oracle.form_like(f), b, nodes, par, u, coefs, load)
    e1, e2 = np.max(np.abs(eta_f - ref) / ref), np.max(np.abs(eta_e - ref) / ref)
    print(f"eta2: kept factor vs oracle {e1:.2e}, re-assembled vs oracle {e2:.2e}")
    assert e1 < 1e-11 and e2 < 1e-11
    # the kept factor's first rows are the Cholesky factor of S G S
    outG, _ = A.local_stiffness_and_rhs(0, dn, dp, dc, load, want_G=True)
    N = A.info[0].N
    L = out["factor"][:, :N, :].cpu().numpy()
    s = out["factor_scale"].cpu().numpy()
    G = outG["G"].cpu().numpy()
    SGS = s[:, :, None] * G * s[:, None, :]
    assert np.all(np.triu(L, 1) == 0.0)
    assert rel(L @ np.transpose(L, (0, 2, 1)), SGS) < 1e-13
    A.close()


def test_energy_error_factored_vanishes_for_exact_linear_solution():
    from helpers import linear_trial_coefficients, poisson2d

    f, b, nodes, par, _, _ = poisson2d(4, 4, k=2, kt=3)
    u = linear_trial_coefficients(nodes, par, 2, (0.2, 1.1, -0.4))
    A = dpg.LocalAssembler(f, [b])
    from camellia2_amd import capi

    out, nbad = A.local_stiffness_and_rhs(0, dev(nodes), dev(par), None, capi.LoadDesc(), want_factor=True)
    assert nbad == 0
    eta2 = A.energy_error_factored(0, out["factor"], dev(u)).cpu().numpy()
    assert np.max(np.abs(eta2)) < 1e-20
    A.close()


def test_autograd_over_K_and_rhs():
    f, b, nodes, par, coefs, load, u, Kb, rb = case(4)
    A = dpg.LocalAssembler(f, [b])
    dn, dp = dev(nodes), dev(par)
    c = dev(coefs).requires_grad_(True)
    K, rhs = dpg.LocalAssemblyFunction.apply(A, 0, dn, dp, c, load)
    ((K * dev(Kb)).sum() + (rhs * dev(rb)).sum()).backward()
    g = A.adjoint(0, dn, dp, dev(coefs), load, Kbar=dev(Kb), rhsbar=dev(rb))
    assert torch.equal(c.grad, g)
    A.close()
