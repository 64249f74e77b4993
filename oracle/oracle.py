"""TEST INFRASTRUCTURE ONLY: ctypes binding of oracle/liboracle.so.

The oracle is the FP64 CPU restatement of the reference's DPG local assembly
(see dpg_oracle.cpp for the reference file:line of every part). Only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py use it, and only
as the parity reference / CPU baseline.
"""
import ctypes as ct
import os

import numpy as np

from camellia2_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ct.CDLL(LIB_PATH)
        P = ct.POINTER
        vp = ct.c_void_p
        L.oracle_factored_cholesky_solve.argtypes = [ct.c_int, ct.c_int, vp, vp, vp, vp, vp]
        L.oracle_schur.argtypes = [ct.c_int, ct.c_int, vp, vp, vp, vp]
        L.oracle_schur_extended.argtypes = [ct.c_int, ct.c_int, vp, vp, vp, vp]
        L.oracle_factored_solve_extended.argtypes = [ct.c_int, ct.c_int, vp, vp, vp, vp, vp]
        L.oracle_sizes.argtypes = [P(capi.FormDesc), P(capi.BucketDesc), vp]
        L.oracle_gauss_rule.argtypes = [ct.c_int, vp, vp]
        L.oracle_gll_nodes.argtypes = [ct.c_int, vp]
        L.oracle_tri_nodes.argtypes = [ct.c_int, vp]
        L.oracle_set_tri_test_basis.argtypes = [ct.c_int]
        L.oracle_set_extended_accumulation.argtypes = [ct.c_int]
        L.oracle_tri_rule.argtypes = [ct.c_int, vp, vp]
        L.oracle_assemble.argtypes = [P(capi.FormDesc), P(capi.BucketDesc), ct.c_int, vp, vp, vp,
                                      P(capi.LoadDesc), vp, vp, vp, vp, vp, vp, vp, vp, ct.c_int]
        L.oracle_energy_error.argtypes = [P(capi.FormDesc), P(capi.BucketDesc), ct.c_int, vp, vp, vp,
                                          P(capi.LoadDesc), vp, vp, ct.c_int]
        L.oracle_graph_norm.argtypes = [P(capi.FormDesc)]
        L.oracle_form_poisson_uw.argtypes = [ct.c_int, P(capi.FormDesc)]
        L.oracle_form_cdr_uw.argtypes = [ct.c_int, P(capi.FormDesc)]
        L.oracle_form_stokes_vgp_uw.argtypes = [ct.c_int, ct.c_double, P(capi.FormDesc)]
        L.oracle_form_convection_diffusion_uw.argtypes = [ct.c_int, P(capi.FormDesc)]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def factored_cholesky_solve(G_colmajor, B_colmajor, l, N, M):
    G = np.array(G_colmajor, dtype=np.float64).copy()
    B = np.array(B_colmajor, dtype=np.float64).copy()
    ll = np.array(l, dtype=np.float64).copy()
    K = np.zeros((M, M))
    rhs = np.zeros(M)
    info = lib().oracle_factored_cholesky_solve(N, M, _p(G), _p(B), _p(ll), _p(K), _p(rhs))
    return K, rhs, info


def sizes(form, bucket):
    out = np.zeros(6, dtype=np.int32)
    rc = lib().oracle_sizes(ct.byref(form), ct.byref(bucket), _p(out))
    assert rc == 0
    return dict(N=int(out[0]), M=int(out[1]), F=int(out[2]), Q1=int(out[3]), Qs1=int(out[4]),
                QG1=int(out[5]))


def gauss_rule(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().oracle_gauss_rule(n, _p(x), _p(w))
    return x, w


def gll_nodes(n):
    x = np.zeros(n + 1)
    lib().oracle_gll_nodes(n, _p(x))
    return x


class tri_test_basis:
    """context manager: the oracle's triangle test-basis variant (0 default,
    1 orthonormal by degree with a centroid RT complement)"""

    def __init__(self, variant):
        self.variant = variant

    def __enter__(self):
        if lib().oracle_set_tri_test_basis(self.variant) != 0:
            raise ValueError(self.variant)

    def __exit__(self, *a):
        lib().oracle_set_tri_test_basis(0)


def tri_nodes(n):
    """warp-and-blend nodes of degree n on the unit triangle, ((n+1)(n+2)/2, 2)"""
    xy = np.zeros(((n + 1) * (n + 2) // 2, 2))
    lib().oracle_tri_nodes(n, _p(xy))
    return xy


def tri_rule(n):
    """collapsed Gauss rule, n*n points (x, y) and weights on the unit triangle"""
    xy, w = np.zeros((n * n, 2)), np.zeros(n * n)
    lib().oracle_tri_rule(n, _p(xy), _p(w))
    return xy, w


def assemble(form, bucket, nodes, parities, coefs=None, load=None, want=("K", "rhs"), nthreads=0):
    """Oracle local assembly of C cells; returns a dict of the requested arrays
    among B (C,M,N), G (C,N,N), l (C,N), K (C,M,M), rhs (C,M), S (C,T,T), g (C,T), info (C)."""
    s = sizes(form, bucket)
    N, M, F = s["N"], s["M"], s["F"]
    T = M - F
    C = nodes.shape[0]
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    parities = np.ascontiguousarray(parities, dtype=np.int8)
    if coefs is not None:
        coefs = np.ascontiguousarray(coefs, dtype=np.float64)
    shapes = dict(B=(C, M, N), G=(C, N, N), l=(C, N), K=(C, M, M), rhs=(C, M), S=(C, T, T), g=(C, T))
    out = {k: np.zeros(shapes[k]) for k in want if k in shapes}
    info = np.zeros(C, dtype=np.int32)
    if load is None:
        load = capi.LoadDesc()
    nbad = lib().oracle_assemble(ct.byref(form), ct.byref(bucket), C, _p(nodes), _p(parities), _p(coefs),
                                 ct.byref(load), _p(out.get("B")), _p(out.get("G")), _p(out.get("l")),
                                 _p(out.get("K")), _p(out.get("rhs")), _p(out.get("S")), _p(out.get("g")),
                                 _p(info), nthreads)
    assert nbad >= 0, "oracle_assemble rejected its arguments"
    out["info"] = info
    return out


def energy_error(form, bucket, nodes, parities, u, coefs=None, load=None, nthreads=0):
    C = nodes.shape[0]
    eta2 = np.zeros(C)
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    parities = np.ascontiguousarray(parities, dtype=np.int8)
    u = np.ascontiguousarray(u, dtype=np.float64)
    if coefs is not None:
        coefs = np.ascontiguousarray(coefs, dtype=np.float64)
    if load is None:
        load = capi.LoadDesc()
    rc = lib().oracle_energy_error(ct.byref(form), ct.byref(bucket), C, _p(nodes), _p(parities), _p(coefs),
                                   ct.byref(load), _p(u), _p(eta2), nthreads)
    assert rc >= 0
    return eta2


def form(name, dim, mu=1.0):
    f = capi.FormDesc()
    if name == "poisson":
        rc = lib().oracle_form_poisson_uw(dim, ct.byref(f))
    elif name == "cdr":
        rc = lib().oracle_form_cdr_uw(dim, ct.byref(f))
    elif name == "stokes":
        rc = lib().oracle_form_stokes_vgp_uw(dim, mu, ct.byref(f))
    elif name == "cd":
        rc = lib().oracle_form_convection_diffusion_uw(dim, ct.byref(f))
    else:
        raise ValueError(name)
    assert rc == 0
    return f


def graph_norm(f):
    rc = lib().oracle_graph_norm(ct.byref(f))
    assert rc == 0
    return f


def form_like(f):
    """A copy of a form descriptor for the oracle (term tables are plain data)."""
    g = capi.FormDesc()
    ct.pointer(g)[0] = f
    return g


def extended_K(G, B, l):
    """K = B^T G^-1 B and rhs = B^T G^-1 l in 80-bit arithmetic from FP64 G (N,N), B (M,N), l (N)."""
    N = G.shape[0]
    M = B.shape[0]
    G = np.ascontiguousarray(G, dtype=np.float64)  # symmetric: row- and column-major agree
    B = np.ascontiguousarray(B, dtype=np.float64)
    l = np.ascontiguousarray(l, dtype=np.float64)
    K = np.zeros((M, M))
    rhs = np.zeros(M)
    lib().oracle_factored_solve_extended(N, M, _p(G), _p(B), _p(l), _p(K), _p(rhs))
    return K, rhs


def extended_schur(K, rhs, F):
    M = K.shape[0]
    T = M - F
    S = np.zeros((T, T))
    g = np.zeros(T)
    K = np.ascontiguousarray(K, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    lib().oracle_schur_extended(M, F, _p(K), _p(rhs), _p(S), _p(g))
    return S, g


def eta2_extended(form, bucket, nodes, parities, u, coefs=None, load=None):
    """eta2 = r^T G^-1 r per cell with G, B, l summed in 80-bit from the same
    FP64 point values, r = l - B u formed in 80-bit (rounded once to FP64) and
    the solve in 80-bit: the reference's energy error evaluated more accurately
    than FP64 reaches, for measuring the FP64 outputs against."""
    lib().oracle_set_extended_accumulation(1)
    try:
        ox = assemble(form, bucket, nodes, parities, coefs, load, want=("B", "G", "l"))
    finally:
        lib().oracle_set_extended_accumulation(0)
    out = np.zeros(nodes.shape[0])
    for c in range(nodes.shape[0]):
        r = (ox["l"][c].astype(np.longdouble) -
             ox["B"][c].T.astype(np.longdouble) @ np.asarray(u[c]).astype(np.longdouble)).astype(np.float64)
        Kr, _ = extended_K(ox["G"][c], r[None, :], ox["l"][c])
        out[c] = Kr[0, 0]
    return out


def with_extended(form, bucket, nodes, parities, coefs=None, load=None, schur=False):
    """Oracle outputs plus extended-precision K, rhs (and S, g): G, B, l summed
    in 80-bit from the same FP64 point values (basis values, weights, Jacobians)
    and then factored, solved and condensed in 80-bit. The distance of the FP64
    outputs from these is the accuracy the reference algorithm reaches in FP64
    on these inputs, end to end."""
    want = ("B", "G", "l", "K", "rhs") + (("S", "g") if schur else ())
    o = assemble(form, bucket, nodes, parities, coefs, load, want=want)
    lib().oracle_set_extended_accumulation(1)
    try:
        ox = assemble(form, bucket, nodes, parities, coefs, load, want=("B", "G", "l"))
    finally:
        lib().oracle_set_extended_accumulation(0)
    F = sizes(form, bucket)["F"]
    x = {k: np.zeros_like(o[k]) for k in ("K", "rhs") + (("S", "g") if schur else ())}
    for c in range(nodes.shape[0]):
        x["K"][c], x["rhs"][c] = extended_K(ox["G"][c], ox["B"][c], ox["l"][c])
        if schur:
            x["S"][c], x["g"][c] = extended_schur(x["K"][c], x["rhs"][c], F)
    return o, x

