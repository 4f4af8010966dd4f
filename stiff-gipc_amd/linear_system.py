"""Python host mirror of the reference's linear-solve call surface.

Class and method names follow StiffGIPC's C++ interface so that callers and
tests read like the reference's own code:

* ``Spmv.srbk_spmv``            -- StiffGIPC/linear_system/utils/spmv.h:19-26
* ``Converter.srbk_convert``    -- StiffGIPC/linear_system/utils/converter.h:40-43
* ``MASPreconditioner``         -- StiffGIPC/MASPreconditioner.cuh:59-107
* ``PCGSolver.solve``           -- StiffGIPC/linear_system/solver/pcg_solver.h:18-45
* ``GlobalLinearSystem.solve_linear_system``
                                -- StiffGIPC/linear_system/linear_system/global_linear_system.h:66

Every call goes through the C ABI into the gfx950 library; arrays are device
arrays (:class:`DevArray`).  Errors raise :class:`SgipcError` carrying the
ABI status (the reference aborts instead).
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _lib
from ._lib import (FP32_IR, FP32_MAS, FP64, PRECISIONS, STOP_REF_RZ, STOP_TRUE_RESIDUAL, Context, DevArray,
                   PcgOpts, PcgStats, SGIPC_CAPACITY, SgipcError, SolveTiming, SceneDesc, check, lib)

__all__ = ["Context", "DevArray", "Spmv", "SpMat", "Converter", "ContactDetector", "MASPreconditioner", "PCGSolver", "GlobalLinearSystem",
           "LinearSystem", "SolveResult", "SgipcError", "FP64", "FP32_MAS", "FP32_IR", "STOP_REF_RZ",
           "STOP_TRUE_RESIDUAL", "pcg_opts", "upload_scene"]


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ptr)


class Spmv:
    def __init__(self, ctx: Context):
        self.ctx = ctx

    def srbk_spmv(self, a, vals, rows, cols, nnz, x, b, y, n_rows):
        """y = a*A*x + b*y, A symmetric stored as unique upper BCOO."""
        if vals.dtype == np.float32:
            return check(lib().sgipc_spmv_srbk_f32(self.ctx.h, a, _p(vals), _p(rows), _p(cols), nnz, _p(x), b, _p(y),
                                                   n_rows), self.ctx.h)
        return check(lib().sgipc_spmv_srbk(self.ctx.h, a, _p(vals), _p(rows), _p(cols), nnz, _p(x), b, _p(y), n_rows),
                     self.ctx.h)


class SpMat:
    """Prepared symmetric matrix (row-tiled layout) for repeated SpMVs."""

    def __init__(self, ctx: Context, vals, rows, cols, nnz, n_rows, want_f32=False):
        self.ctx = ctx
        h = ctypes.c_void_p()
        check(lib().sgipc_spmat_create(ctx.h, ctypes.byref(h), _p(vals), _p(rows), _p(cols), nnz, n_rows,
                                       int(want_f32)), ctx.h)
        self.h = h

    def apply(self, a, x, b, y):
        """y = a*A*x + b*y on device arrays (float64, or float32 with want_f32)."""
        if x.dtype == np.float32:
            return check(lib().sgipc_spmat_apply_f32(self.h, a, _p(x), b, _p(y)), self.ctx.h)
        return check(lib().sgipc_spmat_apply(self.h, a, _p(x), b, _p(y)), self.ctx.h)

    def apply_f32_values(self, a, x, b, y):
        """y = a*A*x + b*y with the FP32 copy of the values and float64 vectors (want_f32)."""
        return check(lib().sgipc_spmat_apply_f32_values(self.h, a, _p(x), b, _p(y)), self.ctx.h)

    def layout(self):
        """(halo_form, halo_max, tl_slots) of the row-tiled layout."""
        v = [ctypes.c_int() for _ in range(3)]
        check(lib().sgipc_spmat_layout(self.h, *[ctypes.byref(t) for t in v]), self.ctx.h)
        return tuple(t.value for t in v)

    def long_rows(self):
        """Rows kept outside the tiles (multiplied by their own kernel)."""
        v = ctypes.c_int()
        check(lib().sgipc_spmat_long_rows(self.h, ctypes.byref(v)), self.ctx.h)
        return v.value

    def close(self):
        if self.h:
            lib().sgipc_spmat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Converter:
    def __init__(self, ctx: Context):
        self.ctx = ctx

    def srbk_convert(self, tvals, trows, tcols, n_triplets, capacity=None):
        """Triplets -> unique upper BCOO; returns (vals, rows, cols, nnz) device arrays."""
        cap = n_triplets if capacity is None else capacity
        ov = self.ctx.empty((max(cap, 1), 9), np.float64)
        orr = self.ctx.empty(max(cap, 1), np.int32)
        oc = self.ctx.empty(max(cap, 1), np.int32)
        nnz = ctypes.c_int()
        check(lib().sgipc_convert_srbk(self.ctx.h, _p(tvals), _p(trows), _p(tcols), n_triplets, _p(ov), _p(orr),
                                       _p(oc), cap, ctypes.byref(nnz)), self.ctx.h)
        return ov, orr, oc, nnz.value


class ContactDetector:
    """lbvh_f / lbvh_e self queries (mlbvh.cu): contact pairs of a surface."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def detect(self, x, rest, faces, n_faces, edges, n_edges, surf, n_surf, n_verts, dhat_sq, body_id=None,
               btype=None, capacity=None):
        """Returns (pairs (device int4), mat_index (device int), counts[5] host)."""
        cap = capacity if capacity is not None else max(64, 8 * (n_surf + n_edges))
        while True:
            pairs = self.ctx.empty((max(cap, 1), 4), np.int32)
            mi = self.ctx.empty(max(cap, 1), np.int32)
            counts = (ctypes.c_int * 5)()
            st = lib().sgipc_contact_detect(self.ctx.h, _p(x), _p(rest), n_verts, _p(faces), n_faces, _p(edges),
                                            n_edges, _p(surf), n_surf, _p(body_id), _p(btype), float(dhat_sq),
                                            _p(pairs), _p(mi), cap, counts)
            if st == SGIPC_CAPACITY and capacity is None:
                cap = counts[0]
                continue
            check(st, self.ctx.h)
            return pairs, mi, np.array(list(counts), dtype=np.int32)

    def barrier(self, x, pairs, mat_index, counts, dhat_sq, kappa, n_verts, hess_offset=0, want_grad=True):
        """Barrier blocks of every pair type; returns (vals, rows, cols, grad) device arrays."""
        n = int(counts[2] + counts[3] + counts[4])
        nt = 16 * int(counts[4]) + 9 * int(counts[3]) + 4 * int(counts[2])
        tv = self.ctx.empty((max(nt, 1), 9), np.float64)
        tr = self.ctx.empty(max(nt, 1), np.int32)
        tc = self.ctx.empty(max(nt, 1), np.int32)
        g = None
        if want_grad:
            g = self.ctx.to_device(np.zeros(3 * n_verts))
        c = (ctypes.c_int * 5)(*[int(v) for v in counts])
        check(lib().sgipc_barrier_gradient_hessian(self.ctx.h, _p(x), _p(pairs), _p(mat_index), n, c, float(dhat_sq),
                                                   float(kappa), hess_offset, _p(g), _p(tv), _p(tr), _p(tc)),
              self.ctx.h)
        return tv, tr, tc, g


class MASPreconditioner:
    """Multilevel additive Schwarz preconditioner on the device."""

    def __init__(self, ctx: Context, bank_size=32):
        self.ctx = ctx
        self.bank_size = bank_size
        h = ctypes.c_void_p()
        check(lib().sgipc_mas_create(ctx.h, ctypes.byref(h), bank_size), ctx.h)
        self.h = h
        self.vert_num = 0

    def set_coarse_mode(self, mode):
        """0: additive levels (MASPreconditioner::preconditioning); 1: a
        symmetric V(1,1) cycle on level 1; 2: that cycle as the coarse solve
        of a symmetric two-grid cycle on level 0.  Applies at the next build."""
        check(lib().sgipc_mas_set_coarse_mode(self.h, int(mode)), self.ctx.h)

    def set_smoothing_weight(self, omega):
        """Weight of the level-0 smoothing in coarse mode 2 (default 0.4)."""
        check(lib().sgipc_mas_set_smoothing_weight(self.h, float(omega)), self.ctx.h)

    def set_level0_half(self, on):
        """FP16 (unit-diagonal scaled) level-0 inverses in the float applies of
        coarse mode 2 (default on).  Applies at the next build."""
        check(lib().sgipc_mas_set_level0_half(self.h, int(bool(on))), self.ctx.h)

    def initPreconditioner_Neighbor(self, vertNum, collision_node_offset, totalNeighborNum, nbr_start, nbr_num,
                                    nbr_list):
        self.vert_num = vertNum
        return check(lib().sgipc_mas_init_neighbor(self.h, vertNum, collision_node_offset, totalNeighborNum,
                                                   _p(nbr_start), _p(nbr_num), _p(nbr_list)), self.ctx.h)

    def initPreconditioner_Matrix(self):
        return check(lib().sgipc_mas_init_matrix(self.h), self.ctx.h)

    def setPreconditioner_bcoo(self, vals, rows, cols, nnz, offset=0, pairs=None, n_pairs=0, precision=FP64):
        return check(lib().sgipc_mas_set_preconditioner_bcoo(self.h, _p(vals), _p(rows), _p(cols), nnz, offset,
                                                             _p(pairs), n_pairs, precision), self.ctx.h)

    def preconditioning(self, R, Z):
        return check(lib().sgipc_mas_preconditioning(self.h, _p(R), _p(Z)), self.ctx.h)

    def levels(self):
        L = ctypes.c_int()
        off = (ctypes.c_int * 8)()
        cnt = (ctypes.c_int * 8)()
        check(lib().sgipc_mas_levels(self.h, ctypes.byref(L), off, cnt), self.ctx.h)
        n = L.value
        return n, np.array(off[:n + 1], np.int32), np.array(cnt[:n], np.int32), int(off[n])

    def tables(self):
        L, off, cnt, total = self.levels()
        ct = self.ctx.empty(max(self.vert_num * (L - 1), 1), np.int32)
        gn = self.ctx.empty(max(total, 1), np.int32)
        check(lib().sgipc_mas_tables(self.h, _p(ct), _p(gn)), self.ctx.h)
        ctab = ct.download()[: self.vert_num * (L - 1)].reshape(self.vert_num, L - 1)
        gnext = gn.download()[:total]
        ct.free()
        gn.free()
        return ctab, gnext

    def inverse_blocks(self):
        L, off, cnt, total = self.levels()
        B = self.bank_size
        nb = B * (B + 1) // 2
        out = self.ctx.empty((total // B, nb, 9), np.float64)
        check(lib().sgipc_mas_inverse_blocks(self.h, _p(out)), self.ctx.h)
        a = out.download()
        out.free()
        return a

    def level2_inverse(self):
        """The dense level-2 inverse of coarse modes 1 and 2 (m x m), or None."""
        m = ctypes.c_int(0)
        check(lib().sgipc_mas_level2_inverse(self.h, None, ctypes.byref(m)), self.ctx.h)
        if m.value == 0:
            return None
        out = self.ctx.empty((m.value, m.value), np.float64)
        check(lib().sgipc_mas_level2_inverse(self.h, _p(out), ctypes.byref(m)), self.ctx.h)
        a = out.download()
        out.free()
        return a

    def close(self):
        if self.h and self.ctx.h:
            lib().sgipc_mas_destroy(self.h)
        self.h = None


def pcg_opts(tol=1e-5, stop_mode=STOP_TRUE_RESIDUAL, max_iter=0, precision=FP64, check_every=8, ir_max_outer=20,
             ir_inner_tol=0.0, fuse=1, flexible=False, ir_replace_every=8, graph=None):
    o = PcgOpts()
    lib().sgipc_pcg_default_opts(ctypes.byref(o))
    o.tol = tol
    o.stop_mode = stop_mode
    o.max_iter = max_iter
    o.precision = PRECISIONS[precision] if isinstance(precision, str) else precision
    o.check_every = check_every
    o.ir_max_outer = ir_max_outer
    o.ir_inner_tol = ir_inner_tol
    o.fuse = int(fuse)
    o.flexible = int(bool(flexible))
    o.ir_replace_every = int(ir_replace_every)
    if graph is not None:
        o.graph = int(bool(graph))
    return o


class PCGSolver:
    def __init__(self, ctx: Context, **opts):
        self.ctx = ctx
        self.opts = pcg_opts(**opts)

    def solve(self, vals, rows, cols, nnz, n_rows, mas, b, x):
        """x = A^-1 b (device arrays); returns the PcgStats of the solve."""
        st = PcgStats()
        status = lib().sgipc_pcg_solve(self.ctx.h, _p(vals), _p(rows), _p(cols), nnz, n_rows,
                                       mas.h if mas is not None else None, _p(b), _p(x), ctypes.byref(self.opts),
                                       ctypes.byref(st))
        check(status, self.ctx.h, allow=(_lib.SGIPC_MAXITER,))
        return st


@dataclasses.dataclass
class SolveResult:
    x: np.ndarray
    iterations: int
    outer_iterations: int
    true_rel_residual: float
    status: int
    timing: dict
    learned: bool = False
    learned_abandoned: bool = False  # a learned run missed twice the plain count; redone without it
    coarse_mode: int = 0             # MAS coarse mode of the run that returned
    level2_exact: bool = False       # its level-1 cycle had the exact dense level-2 solve
    omega: float = 0.0               # mode-2 smoothing weight
    fp64_fallback: int = 0           # 1: a failed mixed-precision solve was redone in FP64; 2: ran FP64 from the start


def upload_scene(ctx: Context, scene):
    """Copy a scene's solver inputs (tets, cloth triangles, hinges, contacts) to
    the device; returns (desc, keepalive)."""
    keep = {}

    def put(name, a, dtype):
        if a is None:
            keep[name] = None
            return None
        d = ctx.to_device(np.ascontiguousarray(a, dtype=dtype))
        keep[name] = d
        return d.ptr

    desc = SceneDesc()
    desc.n_verts = scene.n
    desc.n_tets = scene.n_tets
    desc.n_pairs = scene.n_pairs
    desc.total_neighbor_num = int(scene.nbr_list.size)
    desc.dt = scene.dt
    desc.dhat_sq = scene.dhat ** 2
    desc.kappa = scene.kappa
    desc.x = put("x", scene.x.reshape(-1), np.float64)
    if scene.n_tets:
        desc.tets = put("tets", scene.tets.reshape(-1), np.uint32)
        desc.dm_inv = put("dm_inv", scene.dm_inv.reshape(-1), np.float64)
        desc.vol = put("vol", scene.vol, np.float64)
        desc.mu = put("mu", scene.mu, np.float64)
        desc.lam = put("lam", scene.lam, np.float64)
    desc.n_tris = scene.n_tris
    if scene.n_tris:
        desc.k_stretch, desc.k_shear, desc.strain_rate = scene.k_stretch, scene.k_shear, scene.strain_rate
        desc.tris = put("tris", scene.tris.reshape(-1), np.uint32)
        desc.dm_inv2 = put("dm_inv2", scene.dm_inv2.reshape(-1), np.float64)
        desc.area = put("area", scene.area, np.float64)
    desc.n_hinges = scene.n_hinges
    if scene.n_hinges:
        desc.k_bend = scene.k_bend
        desc.rest_x = put("rest_x", scene.rest.reshape(-1), np.float64)
        desc.hinge_edges = put("hinge_edges", scene.edges.reshape(-1).astype(np.int64) & 0xFFFFFFFF, np.uint32)
        desc.hinge_adj = put("hinge_adj", scene.adj.reshape(-1).astype(np.int64) & 0xFFFFFFFF, np.uint32)
    desc.mass = put("mass", scene.mass, np.float64)
    desc.pairs = put("pairs", scene.pairs.reshape(-1) if scene.n_pairs else None, np.int32)
    desc.pair_x = put("pair_x", scene.pair_x.reshape(-1) if scene.pair_x is not None and scene.n_pairs else None,
                      np.float64)
    nb = getattr(scene, "n_bodies", 0)
    desc.n_bodies = nb
    if nb:
        desc.n_abd_points = scene.n_abd_points
        desc.kappa_abd = scene.kappa_abd
        desc.q = put("q", scene.q.reshape(-1), np.float64)
        desc.q_tilde = put("q_tilde", scene.q_tilde.reshape(-1), np.float64)
        desc.mass10 = put("mass10", scene.mass10.reshape(-1), np.float64)
        desc.body_volume = put("body_volume", scene.body_volume, np.float64)
        desc.body_fixed = put("body_fixed", scene.body_fixed, np.int32)
        desc.point_body = put("point_body", scene.point_body, np.int32)
        desc.x_bar = put("x_bar", scene.x_bar.reshape(-1), np.float64)
        desc.motor_speed = getattr(scene, "motor_speed", 31.4)
        desc.motor_strength = getattr(scene, "motor_strength", 10.0)
    bt = getattr(scene, "boundary_type", None)
    desc.boundary_type = put("boundary_type", bt, np.int32) if bt is not None else None
    desc.nbr_start = put("nbr_start", scene.nbr_start, np.uint32)
    desc.nbr_num = put("nbr_num", scene.nbr_num, np.uint32)
    desc.nbr_list = put("nbr_list", scene.nbr_list if scene.nbr_list.size else np.zeros(1, np.uint32), np.uint32)
    surf = getattr(scene, "contact_surface", None)
    if surf is not None:
        # contacts detected on every solve from the surface (faces, edges, surface vertices)
        desc.detect_contacts = 1
        desc.n_faces, desc.n_edges, desc.n_surf = surf.faces.shape[0], surf.edges.shape[0], surf.surf.size
        desc.faces = put("c_faces", surf.faces.reshape(-1), np.uint32)
        desc.edges = put("c_edges", surf.edges.reshape(-1), np.uint32)
        desc.surf_verts = put("c_surf", surf.surf, np.uint32)
        desc.body_id = put("c_body", surf.body_id, np.int32) if surf.body_id is not None else None
        if not desc.rest_x:
            desc.rest_x = put("rest_x", scene.rest.reshape(-1), np.float64)
    part = getattr(scene, "part_id", None)
    if part is not None:
        # host array (read at system creation)
        hp = np.ascontiguousarray(part, dtype=np.int32)
        keep["part_id"] = hp
        desc.part_id = hp.ctypes.data
    return desc, keep


class GlobalLinearSystem:
    """Whole Newton linear solve on the device: assembly (contacts, tets,
    bending hinges, cloth triangles, mass), conversion to unique upper BCOO,
    MAS rebuild, PCG."""

    def __init__(self, ctx: Context, scene, bank_size=32):
        self.ctx = ctx
        self.scene = scene
        self.desc, self._keep = upload_scene(ctx, scene)
        h = ctypes.c_void_p()
        check(lib().sgipc_system_create(ctx.h, ctypes.byref(h), ctypes.byref(self.desc), bank_size), ctx.h)
        self.h = h
        self.b = ctx.to_device(scene.rhs, np.float64)
        n_sys = getattr(scene, "n_sys", scene.n)
        self.x = ctx.empty(3 * n_sys, np.float64)
        self.grad = ctx.empty(3 * n_sys, np.float64)

    def solve_linear_system(self, tol=1e-5, precision="fp64", stop_mode=STOP_TRUE_RESIDUAL, want_grad=False,
                            **kw):
        o = pcg_opts(tol=tol, precision=precision, stop_mode=stop_mode, **kw)
        st = PcgStats()
        tm = SolveTiming()
        status = lib().sgipc_system_solve(self.h, self.b.vp, self.x.vp, self.grad.vp if want_grad else None,
                                          ctypes.byref(o), ctypes.byref(st), ctypes.byref(tm))
        check(status, self.ctx.h, allow=(_lib.SGIPC_MAXITER,))
        return st, tm

    def bcoo(self):
        """Host copies of the current unique upper BCOO (vals (nnz,9), rows, cols)."""
        v, r, c, n = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int()
        check(lib().sgipc_system_bcoo(self.h, ctypes.byref(v), ctypes.byref(r), ctypes.byref(c), ctypes.byref(n)),
              self.ctx.h)
        nnz = n.value
        vals = np.empty((nnz, 9))
        rows = np.empty(nnz, np.int32)
        cols = np.empty(nnz, np.int32)
        for host, dev in ((vals, v), (rows, r), (cols, c)):
            if nnz:
                check(lib().sgipc_memcpy_dtoh(self.ctx.h, host.ctypes.data, dev, host.nbytes), self.ctx.h)
        return vals, rows, cols

    def mas_handle(self):
        return lib().sgipc_system_mas(self.h)

    def contacts(self):
        """Host copies of the last solve's contact pairs (n, 4) and MatIndex
        (None unless the scene detects its contacts)."""
        p, m, n = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int()
        check(lib().sgipc_system_contacts(self.h, ctypes.byref(p), ctypes.byref(m), ctypes.byref(n)), self.ctx.h)
        pairs = np.empty((n.value, 4), np.int32)
        if n.value:
            check(lib().sgipc_memcpy_dtoh(self.ctx.h, pairs.ctypes.data, p, pairs.nbytes), self.ctx.h)
        mi = None
        if m.value:
            mi = np.empty(n.value, np.int32)
            if n.value:
                check(lib().sgipc_memcpy_dtoh(self.ctx.h, mi.ctypes.data, m, mi.nbytes), self.ctx.h)
        return pairs, mi

    def close(self):
        if self.h and self.ctx.h:
            lib().sgipc_system_destroy(self.h)
        self.h = None


class LinearSystem:
    """Convenience owner of a Context + GlobalLinearSystem for one scene."""

    def __init__(self, scene, bank_size=32, device=0, coarse_mode=0):
        self.ctx = Context(device)
        self.sys = GlobalLinearSystem(self.ctx, scene, bank_size)
        self.scene = scene
        self.set_coarse_mode(coarse_mode)

    def set_coarse_mode(self, mode):
        """0: additive MAS levels (the reference); 1: V(1,1) cycle on level 1;
        2: two-grid cycle on level 0 around it."""
        check(lib().sgipc_mas_set_coarse_mode(self.sys.mas_handle(), int(mode)), self.ctx.h)

    def set_smoothing_weight(self, omega):
        """Weight of the level-0 smoothing in coarse mode 2 (default 0.4)."""
        check(lib().sgipc_mas_set_smoothing_weight(self.sys.mas_handle(), float(omega)), self.ctx.h)

    def set_level0_half(self, on):
        """FP16 (unit-diagonal scaled) level-0 inverses in the float applies of
        coarse mode 2 (default on)."""
        check(lib().sgipc_mas_set_level0_half(self.sys.mas_handle(), int(bool(on))), self.ctx.h)

    def solve(self, tol=1e-5, precision="fp64", stop_mode=STOP_TRUE_RESIDUAL, want_grad=False, **kw):
        st, tm = self.sys.solve_linear_system(tol=tol, precision=precision, stop_mode=stop_mode, want_grad=want_grad,
                                              **kw)
        x = self.sys.x.download()
        timing = {k: getattr(tm, k) for k, _ in SolveTiming._fields_}
        return SolveResult(x=x, iterations=st.iterations, outer_iterations=st.outer_iterations,
                           true_rel_residual=st.rel_residual, status=st.status, timing=timing,
                           learned=st.learned == 1, learned_abandoned=st.learned == 2,
                           coarse_mode=st.coarse_mode, level2_exact=st.level2_exact == 1, omega=st.omega,
                           fp64_fallback=int(st.fp64_fallback))

    def collect(self, stride=1, max_samples=256, tol=1e-5, precision="fp64", b=None):
        """Training samples (MpnnData) from one plain solve of the scene's rhs
        (or of b, host array): residuals every `stride` iterations."""
        from .learned import MpnnData
        o = pcg_opts(tol=tol, precision=precision)
        rhs = self.sys.b if b is None else self.ctx.to_device(np.ascontiguousarray(b, np.float64))
        h = ctypes.c_void_p()
        check(lib().sgipc_system_collect(self.sys.h, rhs.vp, ctypes.byref(o), int(stride), int(max_samples),
                                         ctypes.byref(h)), self.ctx.h)
        return MpnnData(self.ctx, h)

    def set_learned(self, net, mode=2):
        """Attach a learned coarse level: mode 0 off, 1 always, 2 automatic."""
        check(lib().sgipc_system_set_learned(self.sys.h, net.h if net is not None else None, int(mode)), self.ctx.h)

    def gradient(self):
        return self.sys.grad.download()

    def close(self):
        self.sys.close()
        self.ctx.close()
