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
    ref = oracle.energy_error(oracle.form_like(f), b, nodes, par, u, coefs, load)
    e1, e2 = np.max(np.abs(eta2 - ref) / ref), np.max(np.abs(eta2f - ref) / ref)
    print(f"eta2 vs oracle: re-assembled {e1:.2e}, from the kept factor {e2:.2e}")
    assert e1 < 1e-11 and e2 < 1e-11


def test_c5_adjoint_against_central_differences(c5):
    f, b, nodes, par, coefs, load, u, A = c5
    dn, dp, dc, du = dev(nodes), dev(par), dev(coefs), dev(u)
    eta2, grad = A.energy_error_and_gradient(0, dn, dp, du, dc, load)
    out, nbad = A.local_stiffness_and_rhs(0, dn, dp, dc, load, want_factor=True)
    assert nbad == 0
    grad_kept = A.adjoint(0, dn, dp, dc, load, u=du, factor=out["factor"], factor_scale=out["factor_scale"])
    grad, grad_kept = grad.cpu().numpy(), grad_kept.cpu().numpy()
    assert np.linalg.norm(grad_kept - grad) <= 1e-12 * np.linalg.norm(grad)

    def eta(c):
        return A.energy_error(0, dn, dp, du, dev(c), load).cpu().numpy()

    h = 1e-6
    for k in range(coefs.shape[1]):
        cp, cm = coefs.copy(), coefs.copy()
        cp[:, k] += h
        cm[:, k] -= h
        fd = (eta(cp) - eta(cm)) / (2 * h)

        def central(r):
            hh = r * (coefs[:, k] if k == 0 else np.ones(coefs.shape[0]))
            cp2, cm2 = coefs.copy(), coefs.copy()
            cp2[:, k] += hh
            cm2[:, k] -= hh
            return (eta(cp2) - eta(cm2)) / (2 * hh)

        rich = (4.0 * central(5e-3) - central(1e-2)) / 3.0
        e_fd = np.linalg.norm(grad[:, k] - fd) / np.linalg.norm(fd)
        e_rich = np.linalg.norm(grad[:, k] - rich) / np.linalg.norm(rich)
        e_fd_rich = np.linalg.norm(fd - rich) / np.linalg.norm(rich)
        print(f"coef {k}: adjoint vs central differences (h=1e-6) {e_fd:.2e}, vs Richardson {e_rich:.2e} "
              f"(the step-1e-6 quotient itself is {e_fd_rich:.2e} from Richardson)")
        assert e_fd < 1e-6
        assert e_rich < 1e-7


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

    for k in range(coefs.shape[1]):
        def central(hh):
            cp, cm = coefs.copy(), coefs.copy()
            cp[:, k] += hh
            cm[:, k] -= hh
            return (loss(cp) - loss(cm)) / (2 * hh)

        # step 1e-6; the Richardson reference with steps relative to eps for slot 0
        fd = central(1e-6)
        rel_h = coefs[:, k] if k == 0 else 1.0
        rich = (4.0 * central(5e-4 * rel_h) - central(1e-3 * rel_h)) / 3.0
        e_fd = np.linalg.norm(grad[:, k] - fd) / np.linalg.norm(fd)
        e_rich = np.linalg.norm(grad[:, k] - rich) / np.linalg.norm(rich)
        print(f"coef {k}: adjoint of <Kbar,K> + <rbar,rhs> vs central differences {e_fd:.2e}, vs Richardson "
              f"{e_rich:.2e}")
        assert e_fd < 1e-6
        assert e_rich < 1e-7  # rounding of K over the Richardson steps, as in test_gpu_adjoint
