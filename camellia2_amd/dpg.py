"""Host-side API over the C ABI (libcamellia2_dpg.so).

Names follow the reference's call surface: a form is the compiled BF/IP/RHS
of a formulation (PoissonFormulation, ConvectionDiffusionReactionFormulation,
StokesVGPFormulation with BF::graphNorm), `LocalAssembler.local_stiffness_and_rhs`
is TBF::localStiffnessMatrixAndRHS for a batch of cells of one element type
(plus the CondensedDofInterpreter Schur complement), `energy_error` is the
per-cell indicator of TSolution::rankLocalEnergyError.

Device memory and streams come from torch (ROCm); every computation runs in
the HIP library, there is no Python or CPU fallback.
"""
import ctypes as ct

import numpy as np

from . import capi
from .capi import check, lib


def form(name, dim=2, mu=1.0):
    """Term tables of a reference formulation (ultraweak, graph norm, f*v load)."""
    f = capi.FormDesc()
    L = lib()
    if name == "poisson":
        check(L.dpg_form_poisson_uw(dim, ct.byref(f)))
    elif name == "cdr":
        check(L.dpg_form_cdr_uw(dim, ct.byref(f)))
    elif name == "stokes":
        check(L.dpg_form_stokes_vgp_uw(dim, mu, ct.byref(f)))
    elif name == "cd":
        check(L.dpg_form_convection_diffusion_uw(dim, ct.byref(f)))
    else:
        raise ValueError(f"unknown formulation {name}")
    return f


def graph_norm(f):
    check(lib().dpg_form_graph_norm(ct.byref(f)))
    return f


def bucket(topology, h1_order, test_degree):
    return capi.BucketDesc(topology, h1_order, test_degree, 0)


def sizes(f, b):
    info = capi.BucketInfo()
    check(lib().dpg_bucket_sizes(ct.byref(f), ct.byref(b), ct.byref(info)))
    return info


def mesh_quad(nx, ny, jitter, seed):
    C = nx * ny
    nodes = np.zeros((C, 4, 2))
    par = np.zeros((C, 4), dtype=np.int8)
    check(lib().dpg_mesh_quad(nx, ny, jitter, seed, nodes.ctypes.data, par.ctypes.data))
    return nodes, par


def mesh_hex(nx, ny, nz, jitter, seed):
    C = nx * ny * nz
    nodes = np.zeros((C, 8, 3))
    par = np.zeros((C, 6), dtype=np.int8)
    check(lib().dpg_mesh_hex(nx, ny, nz, jitter, seed, nodes.ctypes.data, par.ctypes.data))
    return nodes, par


def mesh_quad_device(nx, ny, jitter, seed, stream=None):
    """mesh_quad generated on the device (dpg_mesh_quad_device): the same bits, as device tensors."""
    import torch

    C = nx * ny
    nodes = torch.empty((C, 4, 2), dtype=torch.float64, device="cuda:0")
    par = torch.empty((C, 4), dtype=torch.int8, device="cuda:0")
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().dpg_mesh_quad_device(nx, ny, jitter, seed, nodes.data_ptr(), par.data_ptr(), ct.c_void_p(s)))
    return nodes, par


def mesh_hex_device(nx, ny, nz, jitter, seed, stream=None):
    """mesh_hex generated on the device (dpg_mesh_hex_device): the same bits, as device tensors."""
    import torch

    C = nx * ny * nz
    nodes = torch.empty((C, 8, 3), dtype=torch.float64, device="cuda:0")
    par = torch.empty((C, 6), dtype=torch.int8, device="cuda:0")
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().dpg_mesh_hex_device(nx, ny, nz, jitter, seed, nodes.data_ptr(), par.data_ptr(), ct.c_void_p(s)))
    return nodes, par


def mesh_quad_graded(n0, levels, jitter, seed):
    """hp mesh graded toward x = 1: (nodes (C,4,2), parities (C,4), refinement level (C,))."""
    L = lib()
    C = check(L.dpg_mesh_quad_graded(n0, levels, jitter, seed, None, None, None))
    nodes = np.zeros((C, 4, 2))
    par = np.zeros((C, 4), dtype=np.int8)
    lvl = np.zeros((C,), dtype=np.int32)
    check(L.dpg_mesh_quad_graded(n0, levels, jitter, seed, nodes.ctypes.data, par.ctypes.data, lvl.ctypes.data))
    return nodes, par, lvl


def load_sine_modes(dim, ncomp, nmodes, kmax, seed):
    L = capi.LoadDesc()
    check(lib().dpg_load_sine_modes(dim, ncomp, nmodes, kmax, seed, ct.byref(L)))
    return L


def rng_uniform(seed, lo, hi, n):
    out = np.zeros(n)
    check(lib().dpg_rng_uniform(seed, lo, hi, out.ctypes.data, n))
    return out


def rng_normal(seed, n):
    out = np.zeros(n)
    check(lib().dpg_rng_normal(seed, out.ctypes.data, n))
    return out


class LocalAssembler:
    """A plan: device basis/cubature tables of the given element types."""

    def __init__(self, f, buckets, precision=capi.PREC_FP64, refine=0, workspace_bytes=0.0, max_batch_cells=0,
                 profile=False):
        import torch  # device memory only

        self._torch = torch
        if not torch.cuda.is_available():
            raise RuntimeError("LocalAssembler needs a ROCm GPU")
        self.form = f
        self.buckets = list(buckets)
        arr = (capi.BucketDesc * len(self.buckets))(*self.buckets)
        opt = capi.Options(precision, refine, max_batch_cells, 1 if profile else 0, float(workspace_bytes))
        plan = ct.c_void_p()
        check(lib().dpg_plan_create(ct.byref(f), arr, len(self.buckets), ct.byref(opt), ct.byref(plan)))
        self._plan = plan
        self.info = [sizes(f, b) for b in self.buckets]

    def clone(self, workspace_bytes=0.0):
        """A second assembler on the same device tables with its own workspace
        (dpg_plan_clone_workspace): the two may run concurrently on different streams."""
        c = object.__new__(LocalAssembler)
        c._torch = self._torch
        c.form = self.form
        c.buckets = list(self.buckets)
        c.info = list(self.info)
        plan = ct.c_void_p()
        check(lib().dpg_plan_clone_workspace(self._plan, float(workspace_bytes), ct.byref(plan)))
        c._plan = plan
        return c

    def close(self):
        if self._plan:
            lib().dpg_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def workspace_bytes(self):
        n = ct.c_size_t()
        check(lib().dpg_plan_workspace(self._plan, ct.byref(n)))
        return n.value

    def timings(self, reset=True):
        """Per-stage kernel times / algorithmic flops / bytes since the last reset (profile=True plans)."""
        t = capi.Timings()
        check(lib().dpg_plan_timings(self._plan, ct.byref(t), 1 if reset else 0))
        out = {}
        for i, name in enumerate(capi.STAGES):
            if t.launches[i]:
                out[name] = {"ms": t.ms[i], "launches": int(t.launches[i]), "flops": t.flops[i], "bytes": t.bytes[i]}
        return out, int(t.cells)

    def tiles(self, bucket):
        """FP64 contraction choices of a bucket: (gram, bmat, update) tile edges and sumfact (0/1)."""
        t = capi.Tiles()
        check(lib().dpg_plan_get_tiles(self._plan, bucket, ct.byref(t)))
        return (t.gram, t.bmat, t.update, t.sumfact)

    def set_tiles(self, bucket, gram, bmat, update, sumfact=None):
        """sumfact None keeps the current G path; 1 asks for the sum-factorised G (quads, hexes)."""
        if sumfact is None:
            sumfact = self.tiles(bucket)[3]
        check(lib().dpg_plan_set_tiles(self._plan, bucket, ct.byref(capi.Tiles(gram, bmat, update, sumfact))))

    def autotune(self, bucket, nodes, parities, coefs=None, load=None, sample=256, stream=None):
        """Times both tile edges per GEMM role on `sample` cells and keeps the faster, then takes the
        sum-factorised G where it is faster and accurate (dpg_plan_autotune); returns the choice
        (gram, bmat, update, sumfact)."""
        t = self._torch
        b = self._batch(nodes, parities, coefs, load)
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_plan_autotune(self._plan, bucket, ct.byref(b), sample, ct.c_void_p(s)))
        return self.tiles(bucket)

    def _outputs(self, bucket, C, dev, out, want_G, want_schur, want_info, want_factor=False):
        t = self._torch
        info = self.info[bucket]
        o = out if out is not None else {}
        if "K" not in o:
            o["K"] = t.empty((C, info.M, info.M), dtype=t.float64, device=dev)
        if "rhs" not in o:
            o["rhs"] = t.empty((C, info.M), dtype=t.float64, device=dev)
        if want_G and "G" not in o:
            o["G"] = t.empty((C, info.N, info.N), dtype=t.float64, device=dev)
        if want_schur and "schur" not in o:
            o["schur"] = t.empty((C, info.T, info.T), dtype=t.float64, device=dev)
        if want_schur and "schur_rhs" not in o:
            o["schur_rhs"] = t.empty((C, info.T), dtype=t.float64, device=dev)
        if want_info and "info" not in o:
            o["info"] = t.zeros((C,), dtype=t.int32, device=dev)
        if want_factor and "factor" not in o:
            o["factor"] = t.empty((C, info.N + info.M + 1, info.N), dtype=t.float64, device=dev)
        if want_factor and "factor_scale" not in o:
            o["factor_scale"] = t.empty((C, info.N), dtype=t.float64, device=dev)
        outs = capi.Outputs()
        for key in ("K", "rhs", "G", "schur", "schur_rhs", "info", "factor", "factor_scale"):
            setattr(outs, key, o[key].data_ptr() if key in o else None)
        return o, outs

    def _batch(self, nodes, parities, coefs, load):
        t = self._torch
        b = capi.Batch()
        b.ncells = nodes.shape[0]
        b.nodes = nodes.data_ptr()
        b.parities = parities.data_ptr()
        b.coefs = coefs.data_ptr() if coefs is not None else None
        if load is not None:
            b.load = load
        assert nodes.is_cuda and nodes.dtype == t.float64 and nodes.is_contiguous()
        assert parities.is_cuda and parities.dtype == t.int8 and parities.is_contiguous()
        if coefs is not None:
            assert coefs.is_cuda and coefs.dtype == t.float64 and coefs.is_contiguous()
        return b

    def local_stiffness_and_rhs(self, bucket, nodes, parities, coefs=None, load=None, want_G=False,
                                want_schur=False, want_info=True, out=None, stream=None, want_factor=False):
        """K (C,M,M), rhs (C,M) [, G (C,N,N), schur (C,T,T), schur_rhs (C,T),
        factor (C,N+M+1,N) and factor_scale (C,N) kept for energy_error_factored
        and adjoint] on the device.

        `nodes`, `parities`, `coefs` are device tensors. Returns (outputs dict, #non-SPD cells)."""
        t = self._torch
        o, outs = self._outputs(bucket, nodes.shape[0], nodes.device, out, want_G, want_schur, want_info,
                                want_factor)
        b = self._batch(nodes, parities, coefs, load)
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        nbad = check(lib().dpg_assemble(self._plan, bucket, ct.byref(b), ct.byref(outs), ct.c_void_p(s)))
        return o, nbad

    def recover_fields(self, bucket, K, rhs, u, stream=None):
        """In place: the field entries of u (C, M) from its trace/flux entries,
        u_F = K_FF^-1 (rhs_F - K_FT u_T) per cell (dpg_recover_fields); K (C,M,M)
        and rhs (C,M) from local_stiffness_and_rhs. Returns the non-SPD cell count."""
        t = self._torch
        C = u.shape[0]
        for a in (K, rhs, u):
            assert a.is_cuda and a.dtype == t.float64 and a.is_contiguous() and a.shape[0] == C
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        return check(lib().dpg_recover_fields(self._plan, bucket, C, ct.c_void_p(K.data_ptr()),
                                              ct.c_void_p(rhs.data_ptr()), ct.c_void_p(u.data_ptr()),
                                              ct.c_void_p(s)))

    def energy_error(self, bucket, nodes, parities, u, coefs=None, load=None, stream=None):
        """eta2[c] = r^T G^-1 r with r = l - B u (TSolution::rankLocalEnergyError reports sqrt)."""
        t = self._torch
        C = nodes.shape[0]
        eta2 = t.empty((C,), dtype=t.float64, device=nodes.device)
        assert u.is_cuda and u.dtype == t.float64 and u.is_contiguous()
        b = self._batch(nodes, parities, coefs, load)
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_energy_error(self._plan, bucket, ct.byref(b), ct.c_void_p(u.data_ptr()),
                                     ct.c_void_p(eta2.data_ptr()), ct.c_void_p(s)))
        return eta2

    def riesz_representation(self, bucket, nodes, parities, u, coefs=None, load=None, stream=None):
        """e (C, N): the error representation G^-1 (l - B u) of every cell (dpg_riesz_representation)."""
        t = self._torch
        C = nodes.shape[0]
        e = t.empty((C, self.info[bucket].N), dtype=t.float64, device=nodes.device)
        assert u.is_cuda and u.dtype == t.float64 and u.is_contiguous()
        b = self._batch(nodes, parities, coefs, load)
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_riesz_representation(self._plan, bucket, ct.byref(b), ct.c_void_p(u.data_ptr()),
                                             ct.c_void_p(e.data_ptr()), ct.c_void_p(s)))
        return e

    def energy_error_factored(self, bucket, factor, u, stream=None):
        """eta2 (C,) from the factor a local_stiffness_and_rhs(..., want_factor=True)
        kept (dpg_energy_error_factored): no re-assembly, no refactorisation."""
        t = self._torch
        C = u.shape[0]
        for a in (factor, u):
            assert a.is_cuda and a.dtype == t.float64 and a.is_contiguous() and a.shape[0] == C
        eta2 = t.empty((C,), dtype=t.float64, device=u.device)
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_energy_error_factored(self._plan, bucket, C, ct.c_void_p(factor.data_ptr()),
                                              ct.c_void_p(u.data_ptr()), ct.c_void_p(eta2.data_ptr()),
                                              ct.c_void_p(s)))
        return eta2

    def adjoint(self, bucket, nodes, parities, coefs, load=None, u=None, Kbar=None, rhsbar=None, factor=None,
                factor_scale=None, want_eta2=False, stream=None):
        """grad (C, n_coefs) of <Kbar, K> + <rhsbar, rhs> + eta2(u) per cell with
        respect to the cell's coefficients (dpg_adjoint; each part only when its
        input is given). With factor/factor_scale from local_stiffness_and_rhs(...,
        want_factor=True) nothing is re-assembled. Returns grad, or (eta2, grad)
        with want_eta2 (needs u)."""
        t = self._torch
        C = nodes.shape[0]
        grad = t.empty((C, self.form.n_coefs), dtype=t.float64, device=nodes.device)
        eta2 = t.empty((C,), dtype=t.float64, device=nodes.device) if want_eta2 else None
        for a in (u, Kbar, rhsbar, factor, factor_scale):
            if a is not None:
                assert a.is_cuda and a.dtype == t.float64 and a.is_contiguous() and a.shape[0] == C
        b = self._batch(nodes, parities, coefs, load)
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        vp = lambda a: ct.c_void_p(a.data_ptr()) if a is not None else None  # noqa: E731
        check(lib().dpg_adjoint(self._plan, bucket, ct.byref(b), vp(u), vp(Kbar), vp(rhsbar), vp(factor),
                                vp(factor_scale), vp(eta2), vp(grad), ct.c_void_p(s)))
        return (eta2, grad) if want_eta2 else grad

    def energy_error_and_gradient(self, bucket, nodes, parities, u, coefs, load=None, stream=None):
        """(eta2 (C,), d eta2 / d coefs (C, n_coefs)) from the adjoint kernels (dpg_adjoint)."""
        return self.adjoint(bucket, nodes, parities, coefs, load, u=u, want_eta2=True, stream=stream)


class AssemblyGraph:
    """All element types of an hp mesh assembled by one captured hipGraph launch.

    `items` is a list of dicts with keys bucket, nodes, parities and optionally
    coefs, load, want_schur, want_G, want_info, out (device tensors, kept alive
    by this object because the graph holds their addresses)."""

    def __init__(self, assembler, items, lanes=4, stream=None):
        t = assembler._torch
        self.assembler = assembler
        self.items = []
        arr = (capi.WorkItem * len(items))()
        for i, it in enumerate(items):
            nodes = it["nodes"]
            o, outs = assembler._outputs(it["bucket"], nodes.shape[0], nodes.device, it.get("out"),
                                         it.get("want_G", False), it.get("want_schur", False),
                                         it.get("want_info", True))
            arr[i].bucket = it["bucket"]
            arr[i].batch = assembler._batch(nodes, it["parities"], it.get("coefs"), it.get("load"))
            arr[i].out = outs
            self.items.append(dict(it, out=o))
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        g = ct.c_void_p()
        check(lib().dpg_graph_create(assembler._plan, arr, len(items), lanes, ct.c_void_p(s), ct.byref(g)))
        self._graph = g

    def launch(self, stream=None):
        t = self.assembler._torch
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_graph_launch(self._graph, ct.c_void_p(s)))

    def outputs(self):
        return [it["out"] for it in self.items]

    def workspace_bytes(self):
        n = ct.c_size_t()
        check(lib().dpg_graph_workspace(self._graph, ct.byref(n)))
        return n.value

    def close(self):
        if self._graph:
            lib().dpg_graph_destroy(self._graph)
            self._graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EnergyErrorFunction:
    """torch.autograd.Function over the adjoint: sum_c eta2_c as a function of the
    per-cell coefficients (C, n_coefs). Used to check the adjoint against finite
    differences; the forward and backward both run in the HIP library."""

    @staticmethod
    def apply(assembler, bucket, nodes, parities, u, coefs, load=None):
        import torch

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, c):
                eta2, grad = assembler.energy_error_and_gradient(bucket, nodes, parities, u, c.detach().contiguous(),
                                                                 load)
                ctx.save_for_backward(grad)
                return eta2

            @staticmethod
            def backward(ctx, g):
                (grad,) = ctx.saved_tensors
                return g[:, None] * grad

        return _F.apply(coefs)


class LocalAssemblyFunction:
    """torch.autograd.Function over the local assembly: (K, rhs) of every cell
    as functions of the per-cell coefficients (C, n_coefs). The forward keeps
    the factorisation, the backward is dpg_adjoint with Kbar, rhsbar = the
    incoming gradients (no re-assembly). Used to check the adjoint against
    finite differences; forward and backward both run in the HIP library."""

    @staticmethod
    def apply(assembler, bucket, nodes, parities, coefs, load=None):
        import torch

        class _F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, c):
                cd = c.detach().contiguous()
                out, nbad = assembler.local_stiffness_and_rhs(bucket, nodes, parities, cd, load, want_factor=True)
                if nbad:
                    raise RuntimeError(f"{nbad} cells with a non-SPD Gram matrix")
                ctx.save_for_backward(cd, out["factor"], out["factor_scale"])
                return out["K"], out["rhs"]

            @staticmethod
            def backward(ctx, gK, grhs):
                cd, fac, fs = ctx.saved_tensors
                kb = gK.contiguous() if gK is not None else None
                rb = grhs.contiguous() if grhs is not None else None
                return assembler.adjoint(bucket, nodes, parities, cd, load, Kbar=kb, rhsbar=rb, factor=fac,
                                         factor_scale=fs)

        return _F.apply(coefs)


# ---------------------------------------------------------------- condensed global system
def vertex_ids(nodes):
    """(C, V) global vertex ids of a mesh given by per-cell vertex coordinates
    (C, V, d) in which shared vertices carry identical coordinates (as the seeded
    generators produce): ids in lexicographic order of the distinct points."""
    C, V, d = nodes.shape
    _, inv = np.unique(nodes.reshape(C * V, d), axis=0, return_inverse=True)
    return inv.reshape(C, V).astype(np.int32)


def skeleton_numbering(f, bucket, cell_vertices):
    """Global index (C, T) of every trace and flux dof of every cell of a
    conforming mesh, and the number of global dofs (dpg_skeleton_numbering)."""
    cv = np.ascontiguousarray(cell_vertices, dtype=np.int32)
    T = sizes(f, bucket).T
    gdof = np.zeros((cv.shape[0], T), dtype=np.int32)
    n = check(lib().dpg_skeleton_numbering(ct.byref(f), ct.byref(bucket), cv.shape[0], _p32(cv), _p32(gdof)))
    return gdof, n


def skeleton_entities(f, bucket, cell_vertices, nglobal):
    """Mesh entity (vertex, edge, face) of every global dof of the same numbering,
    and the number of entities (dpg_skeleton_entities)."""
    cv = np.ascontiguousarray(cell_vertices, dtype=np.int32)
    ent = np.zeros(int(nglobal), dtype=np.int32)
    n = check(lib().dpg_skeleton_entities(ct.byref(f), ct.byref(bucket), cv.shape[0], _p32(cv), _p32(ent)))
    return ent, n


def skeleton_nodes(f, bucket):
    """Reference coordinates (T, dim) of the nodes of a cell's trace and flux dofs."""
    s = sizes(f, bucket)
    xi = np.zeros((s.T, f.spatial_dim))
    check(lib().dpg_skeleton_nodes(ct.byref(f), ct.byref(bucket), xi.ctypes.data_as(ct.POINTER(ct.c_double))))
    return xi


# cell-local vertices of each side (the reference cells' side node maps)
SIDE_VERTICES = {
    capi.QUAD: [(0, 1), (1, 2), (2, 3), (3, 0)],
    capi.TRI: [(0, 1), (1, 2), (2, 0)],
    capi.HEX: [(0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 4, 7, 3), (0, 3, 2, 1), (4, 5, 6, 7)],
}


def side_blocks(f, bucket):
    """(variable, kind, side, offset within the trailing T dofs, cardinality) of
    every trace and flux block, in the cell's dof order"""
    dim = f.spatial_dim
    nsides = len(SIDE_VERTICES[bucket.topology])
    out, off = [], 0
    for v in range(f.n_trial):
        kind = f.trial_kind[v]
        if kind not in (capi.TRACE, capi.FLUX):
            continue
        card = (bucket.h1_order + 1) ** (dim - 1) if kind == capi.TRACE else bucket.h1_order ** (dim - 1)
        for s in range(nsides):
            out.append((v, kind, s, off, card))
            off += card
    return out


def boundary_trace_mask(f, bucket, gdof, cell_vertices, nglobal):
    """int32 (nglobal): 1 on the global trace dofs of boundary sides (sides that
    belong to one cell only), the dofs a Dirichlet condition on the trace fixes"""
    cv = np.asarray(cell_vertices)
    sides = SIDE_VERTICES[bucket.topology]
    keys = np.sort(cv[:, np.array(sides)], axis=2)  # (C, S, nodes per side)
    flat = keys.reshape(-1, keys.shape[2])
    _, inv, cnt = np.unique(flat, axis=0, return_inverse=True, return_counts=True)
    on_bd = (cnt[inv.ravel()] == 1).reshape(keys.shape[:2])  # (C, S)
    mask = np.zeros(nglobal, dtype=np.int32)
    for v, kind, s, off, card in side_blocks(f, bucket):
        if kind == capi.TRACE:
            cells = np.flatnonzero(on_bd[:, s])
            mask[gdof[cells, off:off + card].ravel()] = 1
    return mask


def _p32(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_int32))


class CondensedSystem:
    """The global condensed (trace and flux) system of a conforming mesh on the
    device: CSR pattern from the cells' global dof indices, and assembly of the
    cells' Schur complements into it (dpg_skeleton_*)."""

    def __init__(self, gdof, nglobal, stream=None):
        import torch

        self._torch = torch
        g = np.ascontiguousarray(gdof, dtype=np.int32)
        self.ncells, self.T = g.shape
        self.nglobal = int(nglobal)
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        h = ct.c_void_p()
        check(lib().dpg_skeleton_create(self.ncells, self.T, _p32(g), self.nglobal, ct.c_void_p(s), ct.byref(h)))
        self._h = h
        nnz = ct.c_int64()
        check(lib().dpg_skeleton_info(h, ct.byref(nnz), None, None, None))
        self.nnz = nnz.value

    def set_blocks(self, entity, nentities, stream=None):
        """Block-Jacobi preconditioning of solve() over the dofs of each entity
        (dpg_skeleton_set_blocks); nentities = 0 returns to point Jacobi."""
        e = np.ascontiguousarray(entity, dtype=np.int32)
        assert e.shape == (self.nglobal,)
        s = stream.cuda_stream if stream is not None else self._torch.cuda.current_stream().cuda_stream
        check(lib().dpg_skeleton_set_blocks(self._h, _p32(e), int(nentities), ct.c_void_p(s)))

    def pattern(self, stream=None):
        """(row_ptr int64 (nglobal+1), col int32 (nnz)) as device tensors (copies)."""
        t = self._torch
        rp = t.empty(self.nglobal + 1, dtype=t.int64, device="cuda")
        col = t.empty(self.nnz, dtype=t.int32, device="cuda")
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_skeleton_pattern(self._h, ct.c_void_p(rp.data_ptr()), ct.c_void_p(col.data_ptr()),
                                         ct.c_void_p(s)))
        return rp, col

    def assemble(self, schur, schur_rhs=None, stream=None):
        """CSR values (nnz) and global load (nglobal) as device tensors."""
        t = self._torch
        assert schur.is_cuda and schur.dtype == t.float64 and schur.is_contiguous()
        assert schur.shape == (self.ncells, self.T, self.T)
        vals = t.empty(self.nnz, dtype=t.float64, device=schur.device)
        rhs = None
        if schur_rhs is not None:
            assert schur_rhs.is_cuda and schur_rhs.is_contiguous() and schur_rhs.shape == (self.ncells, self.T)
            rhs = t.empty(self.nglobal, dtype=t.float64, device=schur.device)
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_skeleton_assemble(self._h, ct.c_void_p(schur.data_ptr()),
                                          ct.c_void_p(schur_rhs.data_ptr() if rhs is not None else 0),
                                          ct.c_void_p(vals.data_ptr()),
                                          ct.c_void_p(rhs.data_ptr() if rhs is not None else 0), ct.c_void_p(s)))
        return vals, rhs

    def solve(self, values, rhs, fixed, x, tol=1e-12, maxit=10000, stream=None):
        """In place on x: the free dofs of values (CSR) x = rhs with x[fixed] kept
        (fixed: int32 0/1 per global dof); Jacobi- (block-Jacobi after set_blocks)
        preconditioned CG on the device.
        Returns (iterations, final relative residual)."""
        t = self._torch
        for a, dt in ((values, t.float64), (rhs, t.float64), (fixed, t.int32), (x, t.float64)):
            assert a.is_cuda and a.dtype == dt and a.is_contiguous()
        it, rel = ct.c_int(), ct.c_double()
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_skeleton_solve(self._h, ct.c_void_p(values.data_ptr()), ct.c_void_p(rhs.data_ptr()),
                                       ct.c_void_p(fixed.data_ptr()), ct.c_void_p(x.data_ptr()), float(tol),
                                       int(maxit), ct.byref(it), ct.byref(rel), ct.c_void_p(s)))
        return it.value, rel.value

    def solve_twolevel(self, coarse, schur, values, rhs, fixed, x, tol=1e-12, maxit=10000, coarse_iters=20,
                       stream=None):
        """solve() preconditioned by two levels (dpg_skeleton_solve_twolevel): the
        damped entity block-Jacobi (or point Jacobi) smoother and a correction on the
        lowest-order level `coarse` (a CoarseLevel of the same mesh), whose
        matrix comes from the cells' Schur blocks `schur` (C, T, T); flexible CG.
        Returns (iterations, final relative residual)."""
        t = self._torch
        for a, dt in ((schur, t.float64), (values, t.float64), (rhs, t.float64), (fixed, t.int32),
                      (x, t.float64)):
            assert a.is_cuda and a.dtype == dt and a.is_contiguous()
        it, rel = ct.c_int(), ct.c_double()
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_skeleton_solve_twolevel(self._h, coarse._h, ct.c_void_p(schur.data_ptr()),
                                                ct.c_void_p(values.data_ptr()), ct.c_void_p(rhs.data_ptr()),
                                                ct.c_void_p(fixed.data_ptr()), ct.c_void_p(x.data_ptr()), float(tol),
                                                int(maxit), int(coarse_iters), ct.byref(it), ct.byref(rel),
                                                ct.c_void_p(s)))
        return it.value, rel.value

    def close(self):
        if self._h:
            lib().dpg_skeleton_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass



def mark_cells(err, rel_threshold, stream=None):
    """flags (int32 device tensor) of the cells with err > rel_threshold * max err,
    and their count (dpg_mark_cells)."""
    import torch

    assert err.is_cuda and err.dtype == torch.float64 and err.is_contiguous()
    flags = torch.empty(err.shape[0], dtype=torch.int32, device=err.device)
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    n = check(lib().dpg_mark_cells(ct.c_void_p(err.data_ptr()), err.shape[0], float(rel_threshold),
                                   ct.c_void_p(flags.data_ptr()), ct.c_void_p(s)))
    return flags, n


# ----------------------------------------------------------- input preparation on the device
def side_parities(topology, cell_vertices, stream=None):
    """(C, S) int8 side parities of a conforming mesh from device cell -> vertex ids (dpg_side_parities)."""
    import torch

    cv = cell_vertices
    assert cv.is_cuda and cv.dtype == torch.int32 and cv.is_contiguous()
    S = len(SIDE_VERTICES[topology])
    par = torch.empty((cv.shape[0], S), dtype=torch.int8, device=cv.device)
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().dpg_side_parities(topology, cv.shape[0], ct.c_void_p(cv.data_ptr()), ct.c_void_p(par.data_ptr()),
                                  ct.c_void_p(s)))
    return par


def sort_cells(type_key, stream=None):
    """Element-type buckets of device int32 type keys: (perm, offsets (nb+1), keys (nb)) (dpg_sort_cells)."""
    import torch

    assert type_key.is_cuda and type_key.dtype == torch.int32 and type_key.is_contiguous()
    C = type_key.shape[0]
    perm = torch.empty(C, dtype=torch.int32, device=type_key.device)
    offsets = torch.empty(C + 1, dtype=torch.int32, device=type_key.device)
    keys = torch.empty(max(C, 1), dtype=torch.int32, device=type_key.device)
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    nb = check(lib().dpg_sort_cells(C, ct.c_void_p(type_key.data_ptr()), ct.c_void_p(perm.data_ptr()),
                                    ct.c_void_p(offsets.data_ptr()), ct.c_void_p(keys.data_ptr()), ct.c_void_p(s)))
    return perm, offsets[:nb + 1], keys[:nb]


def gather_cells(perm, nodes, parities, coefs=None, stream=None):
    """nodes, parities (and coefs) of the cells in perm order, on the device (dpg_gather_cells)."""
    import torch

    C = perm.shape[0]
    out_n = torch.empty_like(nodes)
    out_p = torch.empty_like(parities)
    out_c = torch.empty_like(coefs) if coefs is not None else None
    nd = nodes[0].numel() if C else 0
    S = parities.shape[1]
    nc = coefs.shape[1] if coefs is not None else 0
    s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().dpg_gather_cells(C, ct.c_void_p(perm.data_ptr()), nd, ct.c_void_p(nodes.data_ptr()),
                                 ct.c_void_p(out_n.data_ptr()), S, ct.c_void_p(parities.data_ptr()),
                                 ct.c_void_p(out_p.data_ptr()), nc,
                                 ct.c_void_p(coefs.data_ptr() if coefs is not None else 0),
                                 ct.c_void_p(out_c.data_ptr() if out_c is not None else 0), ct.c_void_p(s)))
    return out_n, out_p, out_c


class CoarseLevel:
    """The lowest-order trace and flux level of a conforming quad or hex mesh
    for two-level preconditioning (dpg_coarse_create): coarse numbering,
    prolongation and the coarse matrix pattern."""

    def __init__(self, f, bucket, cell_vertices, stream=None):
        import torch

        cv = np.ascontiguousarray(cell_vertices, dtype=np.int32)
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        h = ct.c_void_p()
        check(lib().dpg_coarse_create(ct.byref(f), ct.byref(bucket), cv.shape[0], _p32(cv), ct.c_void_p(s),
                                      ct.byref(h)))
        self._h = h
        n, nnz, tc = ct.c_int64(), ct.c_int64(), ct.c_int()
        check(lib().dpg_coarse_info(h, ct.byref(n), ct.byref(nnz), ct.byref(tc)))
        self.nglobal, self.nnz, self.T = n.value, nnz.value, tc.value

    def close(self):
        if self._h:
            lib().dpg_coarse_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _hp_cells(cell_bucket, cell_vertices, vertex_coords):
    cb = np.ascontiguousarray(cell_bucket, dtype=np.int32)
    cv = np.full((cb.shape[0], 4), -1, dtype=np.int32)
    cvi = np.asarray(cell_vertices)
    cv[:, :cvi.shape[1]] = cvi
    X = np.ascontiguousarray(vertex_coords, dtype=np.float64)
    return cb, cv, X


def hp_numbering(f, buckets, cell_bucket, cell_vertices, vertex_coords):
    """The constrained numbering of an hp mesh on the host (dpg_hp_numbering):
    (nglobal, per-cell global dof lists (offsets, dofs), and M as CSR over the
    cells' local trace/flux rows: row_ptr, entry_dof, entry_weight).
    vertex_coords (nvertices, 2) holds the coordinates of the vertex ids."""
    cb, cv, X = _hp_cells(cell_bucket, cell_vertices, vertex_coords)
    arr = (capi.BucketDesc * len(buckets))(*buckets)
    sz = np.zeros(4, dtype=np.int64)
    args = (ct.byref(f), arr, len(buckets), cb.shape[0], _p32(cb), _p32(cv), X.shape[0],
            X.ctypes.data_as(ct.POINTER(ct.c_double)), sz.ctypes.data_as(ct.POINTER(ct.c_int64)))
    check(lib().dpg_hp_numbering(*args, None, None, None, None, None))
    off = np.zeros(cb.shape[0] + 1, dtype=np.int32)
    dofs = np.zeros(int(sz[1]), dtype=np.int32)
    rp = np.zeros(int(sz[2]) + 1, dtype=np.int32)
    ed = np.zeros(int(sz[3]), dtype=np.int32)
    ew = np.zeros(int(sz[3]))
    check(lib().dpg_hp_numbering(*args, _p32(off), _p32(dofs), _p32(rp), _p32(ed),
                                 ew.ctypes.data_as(ct.POINTER(ct.c_double))))
    return int(sz[0]), off, dofs, rp, ed, ew


class HpSystem:
    """The condensed system of a 2D hp mesh (quads and triangles, hanging sides,
    one element type per cell) on the device (dpg_hp_system_*): the minimum
    rule's constrained numbering, assembly of the element types' Schur blocks
    and the local trace/flux coefficients of a global vector."""

    def __init__(self, f, buckets, cell_bucket, cell_vertices, vertex_coords, stream=None):
        import torch

        self._torch = torch
        self.buckets = list(buckets)
        cb, cv, X = _hp_cells(cell_bucket, cell_vertices, vertex_coords)
        self.cell_bucket = cb
        arr = (capi.BucketDesc * len(self.buckets))(*self.buckets)
        s = stream.cuda_stream if stream is not None else torch.cuda.current_stream().cuda_stream
        h = ct.c_void_p()
        check(lib().dpg_hp_system_create(ct.byref(f), arr, len(self.buckets), cb.shape[0], _p32(cb), _p32(cv),
                                         X.shape[0], X.ctypes.data_as(ct.POINTER(ct.c_double)), ct.c_void_p(s),
                                         ct.byref(h)))
        self._h = h
        n, ne, sk = ct.c_int64(), ct.c_int64(), ct.c_void_p()
        check(lib().dpg_hp_system_info(h, ct.byref(n), ct.byref(ne), ct.byref(sk)))
        self.nglobal, self.nentities, self._sk = n.value, ne.value, sk
        nnz = ct.c_int64()
        check(lib().dpg_skeleton_info(sk, ct.byref(nnz), None, None, None))
        self.nnz = nnz.value
        self.sizes = [sizes(f, b) for b in self.buckets]

    def dofs(self):
        """(fixed, entity, rep_cell, rep_local) host int32 arrays of nglobal"""
        out = [np.zeros(self.nglobal, dtype=np.int32) for _ in range(4)]
        check(lib().dpg_hp_system_dofs(self._h, *[_p32(a) for a in out]))
        return tuple(out)

    def assemble(self, schur, schur_rhs, stream=None):
        """CSR values (nnz) and rhs (nglobal) from the per-bucket Schur blocks (lists of device tensors)."""
        t = self._torch
        nb = len(self.buckets)
        sp = (ct.c_void_p * nb)(*[x.data_ptr() if x is not None else None for x in schur])
        gp = (ct.c_void_p * nb)(*[x.data_ptr() if x is not None else None for x in schur_rhs])
        vals = t.empty(self.nnz, dtype=t.float64, device="cuda:0")
        rhs = t.empty(self.nglobal, dtype=t.float64, device="cuda:0")
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_hp_system_assemble(self._h, sp, gp, ct.c_void_p(vals.data_ptr()),
                                           ct.c_void_p(rhs.data_ptr()), ct.c_void_p(s)))
        return vals, rhs

    def local(self, x, u, stream=None):
        """writes the trace/flux entries of every u[b] (C_b, M_b) from the global vector x"""
        t = self._torch
        nb = len(self.buckets)
        up = (ct.c_void_p * nb)(*[a.data_ptr() if a is not None else None for a in u])
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        check(lib().dpg_hp_system_local(self._h, ct.c_void_p(x.data_ptr()), up, ct.c_void_p(s)))

    def solve(self, values, rhs, fixed, x, tol=1e-12, maxit=10000, blocks=True, stream=None):
        """CG on the assembled system (dpg_skeleton_solve on its skeleton), block
        Jacobi over the mesh's vertices and sides unless blocks is False."""
        t = self._torch
        s = stream.cuda_stream if stream is not None else t.cuda.current_stream().cuda_stream
        if blocks:
            _, ent, _, _ = self.dofs()
            check(lib().dpg_skeleton_set_blocks(self._sk, _p32(ent), int(self.nentities), ct.c_void_p(s)))
        it, rel = ct.c_int(), ct.c_double()
        check(lib().dpg_skeleton_solve(self._sk, ct.c_void_p(values.data_ptr()), ct.c_void_p(rhs.data_ptr()),
                                       ct.c_void_p(fixed.data_ptr()), ct.c_void_p(x.data_ptr()), float(tol),
                                       int(maxit), ct.byref(it), ct.byref(rel), ct.c_void_p(s)))
        return it.value, rel.value

    def close(self):
        if self._h:
            lib().dpg_hp_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
