"""Seeded synthetic workloads shared by the benchmark and the tests.

hp convection-diffusion: the graded quad mesh of dpg_mesh_quad_graded with a
per-cell field degree p and test enrichment dp; cells of equal (p, dp) form
one element type (bucket), as Camellia's ElementType keys them
(GlobalDofAssignment.cpp:615-644: trial order p+1, test order p+dp)."""
import numpy as np

from . import capi, dpg

SEED_C3 = 0xC0FFEE + 3
SEED_C1 = 0xC0FFEE + 1


def cd_hp(n0=16, levels=10, seed=SEED_C3, jitter=0.05, pmax=6, dpmax=3):
    """hp-adaptive 2D convection-diffusion (n0=16, levels=10: 49,360 quads).

    Mesh graded toward the outflow edge x = 1; per-cell p ~ U{1..pmax} and
    dp ~ U{1..dpmax} (H1 order p+1, test degree p+dp); per-cell eps
    log-uniform on [1e-4, 1e-2]; beta one unit vector at a seeded angle for
    all cells; load: 8 sine modes with k in 1..6.
    Returns (form, nodes, parities, coefs, load, p, dp, level)."""
    f = dpg.form("cd", 2)
    nodes, par, lvl = dpg.mesh_quad_graded(n0, levels, jitter, seed)
    C = nodes.shape[0]
    p = np.minimum(np.floor(dpg.rng_uniform(seed + 1, 1.0, pmax + 1.0, C)), pmax).astype(np.int32)
    dp = np.minimum(np.floor(dpg.rng_uniform(seed + 2, 1.0, dpmax + 1.0, C)), dpmax).astype(np.int32)
    eps = 10.0 ** dpg.rng_uniform(seed + 3, -4.0, -2.0, C)
    theta = dpg.rng_uniform(seed + 4, 0.0, 2 * np.pi, 1)[0]
    coefs = np.stack([eps, np.full(C, np.cos(theta)), np.full(C, np.sin(theta))], axis=1)
    load = dpg.load_sine_modes(2, 1, 8, 6, SEED_C1 + 100)
    return f, nodes, par, coefs, load, p, dp, lvl


def hp_groups(p, dp):
    """Cells by element type {(h1_order, test_degree): cell indices}, most work first."""
    keys = sorted(set(zip(p.tolist(), dp.tolist())), key=lambda k: (-(k[0] + k[1]), -k[0]))
    return {(a + 1, a + b): np.nonzero((p == a) & (dp == b))[0] for a, b in keys}


def hp_buckets(groups, topology=capi.QUAD):
    return [dpg.bucket(topology, k, kt) for (k, kt) in groups]
