"""Shared helpers: seeded workloads and error norms."""
import numpy as np

from camellia2_amd import capi, dpg

# seeds: 0xC0FFEE + configuration number
SEED = {1: 0xC0FFEE + 1, 2: 0xC0FFEE + 2, 3: 0xC0FFEE + 3, 4: 0xC0FFEE + 4, 5: 0xC0FFEE + 5}


def rel_fro(a, b):
    """relative Frobenius error of a against the reference b, per cell, max over cells"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    axes = tuple(range(1, a.ndim))
    num = np.sqrt(np.sum((a - b) ** 2, axis=axes))
    den = np.sqrt(np.sum(b ** 2, axis=axes))
    return float(np.max(num / np.maximum(den, 1e-300)))


def poisson2d(nx, ny, k=3, kt=4, seed=SEED[1], jitter=0.10):
    """2D Poisson ultraweak, 8-mode sine load, jittered quads (configuration C1 at 16x16)."""
    f = dpg.form("poisson", 2)
    b = dpg.bucket(capi.QUAD, k, kt)
    nodes, par = dpg.mesh_quad(nx, ny, jitter, seed)
    load = dpg.load_sine_modes(2, 1, 8, 6, seed + 100)
    return f, b, nodes, par, None, load


def cdr2d(nx, ny, k=3, kt=4, seed=SEED[5], jitter=0.10):
    """2D convection-diffusion-reaction with per-cell [eps, beta_x, beta_y, c] (configuration C5 at 64x64)."""
    f = dpg.form("cdr", 2)
    b = dpg.bucket(capi.QUAD, k, kt)
    nodes, par = dpg.mesh_quad(nx, ny, jitter, seed)
    C = nx * ny
    eps = 10.0 ** dpg.rng_uniform(seed + 1, -3.0, 0.0, C)
    theta = dpg.rng_uniform(seed + 2, 0.0, 2 * np.pi, 1)[0]
    mag = dpg.rng_uniform(seed + 3, 0.5, 2.0, C)
    react = dpg.rng_uniform(seed + 4, 0.0, 1.0, C)
    coefs = np.stack([eps, mag * np.cos(theta), mag * np.sin(theta), react], axis=1)
    load = dpg.load_sine_modes(2, 1, 8, 6, SEED[1] + 100)
    return f, b, nodes, par, coefs, load


def stokes2d(nx, ny, k=4, kt=5, seed=SEED[2], jitter=0.10):
    """2D Stokes VGP ultraweak, 8 sine modes per force component (configuration C2 at 128x128)."""
    f = dpg.form("stokes", 2, 1.0)
    b = dpg.bucket(capi.QUAD, k, kt)
    nodes, par = dpg.mesh_quad(nx, ny, jitter, seed)
    load = dpg.load_sine_modes(2, 2, 8, 6, seed + 100)
    return f, b, nodes, par, None, load


def poisson3d(nx, ny, nz, k=4, kt=5, seed=SEED[4], jitter=0.05):
    """3D Poisson ultraweak, 16-mode sine load, jittered hexes (configuration C4 at 32^3)."""
    f = dpg.form("poisson", 3)
    b = dpg.bucket(capi.HEX, k, kt)
    nodes, par = dpg.mesh_hex(nx, ny, nz, jitter, seed)
    load = dpg.load_sine_modes(3, 1, 16, 6, seed + 100)
    return f, b, nodes, par, None, load



# the hp workload lives in the package (bench.py uses it too)
from camellia2_amd.workloads import cd_hp, hp_groups  # noqa: E402,F401
