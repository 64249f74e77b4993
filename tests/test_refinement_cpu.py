"""Why the mixed-precision forms offer no FP64 refinement step on G^-1 B
(dpg_options.refine is rejected): refinement removes the error of a
low-precision solve, not the error of a low-precision G. On a cell of the
32^3-hex workload (p = 3, test order 5, the graph-norm Gram matrix of
condition ~2e6 after equilibration), with G, B summed in 80-bit and G rounded
to FP32 as a stand-in for an FP32 contraction:

- K from the rounded G (factored in FP64) is ~6e-3 from the 80-bit K;
- refining X = G^-1 B with residuals against that same rounded G converges to
  its own G^-1 B and leaves the error where it was;
- refining with residuals against the FP64 G does reach the FP64 floor, but
  needs the FP64 G, i.e. the FP64 contraction the mixed forms were to save.

Runs on the oracle (CPU) only."""
import numpy as np
import scipy.linalg as sla

from camellia2_amd import capi, dpg
from oracle import oracle


def test_refinement_cannot_remove_the_rounding_of_G():
    f = dpg.form("poisson", 3)
    b = dpg.bucket(capi.HEX, 4, 5)
    nodes, par = dpg.mesh_hex(32, 32, 32, 0.05, 0xC0FFEE + 4)
    nodes, par = nodes[:1], par[:1]  # one cell, h = 1/32
    ff = oracle.form_like(f)
    oracle.lib().oracle_set_extended_accumulation(1)
    try:
        ox = oracle.assemble(ff, b, nodes, par, None, None, want=("B", "G", "l"))
    finally:
        oracle.lib().oracle_set_extended_accumulation(0)
    G, B, l = ox["G"][0], ox["B"][0].T.copy(), ox["l"][0]  # B: (N, M)
    Kx, _ = oracle.extended_K(G, B.T, l)
    nx = np.linalg.norm(Kx)

    def k_of(X):
        return B.T @ X

    G32 = G.astype(np.float32).astype(np.float64)
    s = 1.0 / np.sqrt(np.diag(G32))
    cho = sla.cho_factor(s[:, None] * G32 * s[None, :], lower=True)

    def solve32(R):  # (G32)^-1 R through the equilibrated FP64 factor of the rounded G
        return s[:, None] * sla.cho_solve(cho, s[:, None] * R)

    X0 = solve32(B)
    e0 = np.linalg.norm(k_of(X0) - Kx) / nx
    Xs, Xd = X0.copy(), X0.copy()
    for _ in range(12):
        Xs = Xs + solve32(B - G32 @ Xs)  # residual against the same rounded G
        Xd = Xd + solve32(B - G @ Xd)    # residual against the FP64 G
    es = np.linalg.norm(k_of(Xs) - Kx) / nx
    ed = np.linalg.norm(k_of(Xd) - Kx) / nx
    print(f"K from the FP32-rounded G {e0:.2e}; refined against it {es:.2e}; refined against the FP64 G {ed:.2e}")
    assert e0 > 1e-6
    assert es > 0.5 * e0
    assert ed < 1e-9
