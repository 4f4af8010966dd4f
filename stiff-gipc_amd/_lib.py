"""ctypes binding of ``include/sgipc.h`` (``lib/libsgipc.so``).

The binding is plumbing for tests, benchmarks and Python callers: every
computation runs in the gfx950 library.  Loading fails loudly when the
library is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsgipc.so")

SGIPC_OK, SGIPC_EMPTY, SGIPC_NOT_SPD, SGIPC_MAXITER, SGIPC_CAPACITY, SGIPC_BAD_INDEX, SGIPC_BAD_ARG, SGIPC_HIP_ERROR = range(8)
STATUS_NAMES = {0: "OK", 1: "EMPTY", 2: "NOT_SPD", 3: "MAXITER", 4: "CAPACITY", 5: "BAD_INDEX", 6: "BAD_ARG",
                7: "HIP_ERROR"}
FP64, FP32_MAS, FP32_IR = 0, 1, 2
STOP_REF_RZ, STOP_TRUE_RESIDUAL = 0, 1
PRECISIONS = {"fp64": FP64, "fp32_mas": FP32_MAS, "fp32_ir": FP32_IR}

vp = ctypes.c_void_p
ci = ctypes.c_int
cd = ctypes.c_double
cf = ctypes.c_float
csz = ctypes.c_size_t


class PcgOpts(ctypes.Structure):
    _fields_ = [("tol", cd), ("stop_mode", ci), ("max_iter", ci), ("precision", ci), ("check_every", ci),
                ("ir_max_outer", ci), ("ir_inner_tol", cd), ("fuse", ci), ("flexible", ci),
                ("ir_replace_every", ci), ("graph", ci)]


class PcgStats(ctypes.Structure):
    _fields_ = [("iterations", ci), ("outer_iterations", ci), ("rel_residual", cd), ("rz", cd), ("rz0", cd),
                ("ms_total", cd), ("status", ci), ("learned", ci), ("coarse_mode", ci), ("level2_exact", ci),
                ("omega", cd), ("fp64_fallback", ci)]


class SceneDesc(ctypes.Structure):
    """sgipc_scene_desc."""
    _fields_ = [("n_verts", ci), ("n_tets", ci), ("n_pairs", ci), ("total_neighbor_num", ci), ("dt", cd),
                ("dhat_sq", cd), ("kappa", cd), ("x", vp), ("tets", vp), ("dm_inv", vp), ("vol", vp), ("mu", vp),
                ("lam", vp), ("mass", vp), ("pairs", vp), ("pair_x", vp), ("nbr_start", vp), ("nbr_num", vp),
                ("nbr_list", vp),
                ("n_tris", ci), ("k_stretch", cd), ("k_shear", cd), ("strain_rate", cd), ("tris", vp),
                ("dm_inv2", vp), ("area", vp),
                ("n_hinges", ci), ("k_bend", cd), ("rest_x", vp), ("hinge_edges", vp), ("hinge_adj", vp),
                ("n_bodies", ci), ("n_abd_points", ci), ("kappa_abd", cd), ("q", vp), ("q_tilde", vp),
                ("mass10", vp), ("body_volume", vp), ("body_fixed", vp), ("point_body", vp), ("x_bar", vp),
                ("boundary_type", vp), ("motor_speed", cd), ("motor_strength", cd),
                ("detect_contacts", ci), ("n_faces", ci), ("n_edges", ci), ("n_surf", ci), ("faces", vp), ("edges", vp),
                ("surf_verts", vp), ("body_id", vp), ("contact_capacity", ci), ("part_id", vp)]


class SolveTiming(ctypes.Structure):
    _fields_ = [("ms_assemble", cd), ("ms_convert", cd), ("ms_mas_build", cd), ("ms_pcg", cd), ("ms_total", cd),
                ("nnz", ci), ("n_triplets", ci)]


_SIGS = {
    "sgipc_abi_version": (ci, []),
    "sgipc_create": (ci, [ctypes.POINTER(vp), ci]),
    "sgipc_destroy": (ci, [vp]),
    "sgipc_set_stream": (ci, [vp, vp]),
    "sgipc_synchronize": (ci, [vp]),
    "sgipc_last_error": (ctypes.c_char_p, [vp]),
    "sgipc_malloc": (ci, [vp, ctypes.POINTER(vp), csz]),
    "sgipc_free": (ci, [vp, vp]),
    "sgipc_memcpy_htod": (ci, [vp, vp, vp, csz]),
    "sgipc_memcpy_dtoh": (ci, [vp, vp, vp, csz]),
    "sgipc_memcpy_dtod": (ci, [vp, vp, vp, csz]),
    "sgipc_memset": (ci, [vp, vp, ci, csz]),
    "sgipc_fem_tet_gradient_hessian": (ci, [vp, vp, vp, vp, vp, vp, vp, ci, cd, ci, vp, vp, vp, vp]),
    "sgipc_barrier_pt_gradient_hessian": (ci, [vp, vp, vp, vp, ci, cd, cd, ci, vp, vp, vp, vp]),
    "sgipc_fem_tri_gradient_hessian": (ci, [vp, vp, vp, vp, vp, ci, cd, cd, cd, cd, ci, vp, vp, vp, vp]),
    "sgipc_quad_bending_prepare": (ci, [vp, vp, vp, vp, ci, vp]),
    "sgipc_quad_bending_gradient_hessian": (ci, [vp, vp, vp, vp, vp, vp, ci, cd, cd, ci, vp, vp, vp, vp]),
    "sgipc_abd_body_gradient_hessian": (ci, [vp, ci, vp, vp, vp, vp, cd, cd, vp, cd, cd, vp, vp, vp, vp]),
    "sgipc_contact_partition": (ci, [vp, vp, vp, vp, ci, ci, vp, vp, ci, vp, vp, vp, vp, vp, ci,
                                     ctypes.POINTER(ci)]),
    "sgipc_abd_jacobi_build": (ci, [vp, vp, vp, vp, ci, ci, vp]),
    "sgipc_abd_jacobi_apply": (ci, [vp, vp, ci, vp, vp]),
    "sgipc_mass_diagonal": (ci, [vp, vp, ci, ci, vp, vp, vp]),
    "sgipc_convert_srbk": (ci, [vp, vp, vp, vp, ci, vp, vp, vp, ci, ctypes.POINTER(ci)]),
    "sgipc_convert_srbk_noalloc": (ci, [vp, vp, vp, vp, ci, vp, vp, vp, ci, ctypes.POINTER(ci)]),
    "sgipc_contact_detect": (ci, [vp, vp, vp, ci, vp, ci, vp, ci, vp, ci, vp, vp, cd, vp, vp, ci, vp]),
    "sgipc_barrier_gradient_hessian": (ci, [vp, vp, vp, vp, ci, vp, cd, cd, ci, vp, vp, vp, vp]),
    "sgipc_friction_prepare": (ci, [vp, vp, vp, ci, cd, cd, vp, vp, vp]),
    "sgipc_ccd_self_step": (ci, [vp, vp, vp, cd, vp, ci, vp, ci, vp, ci, vp, vp, cd, cd, ctypes.POINTER(cd),
                                 ctypes.POINTER(ci)]),
    "sgipc_ground_step": (ci, [vp, vp, vp, ci, vp, cd, vp, cd, ctypes.POINTER(cd)]),
    "sgipc_injective_step": (ci, [vp, vp, vp, ci, vp, cd, cd, ctypes.POINTER(cd)]),
    "sgipc_friction_gradient_hessian": (ci, [vp, vp, vp, vp, vp, ci, vp, vp, vp, vp, cd, cd, ci, vp, vp, vp, vp]),
    "sgipc_ground_friction_prepare": (ci, [vp, vp, vp, ci, vp, cd, cd, cd, vp]),
    "sgipc_ground_friction_gradient_hessian": (ci, [vp, vp, vp, vp, ci, vp, vp, cd, cd, ci, vp, vp, vp, vp]),
    "sgipc_spmv_srbk": (ci, [vp, cd, vp, vp, vp, ci, vp, cd, vp, ci]),
    "sgipc_spmv_srbk_f32": (ci, [vp, cf, vp, vp, vp, ci, vp, cf, vp, ci]),
    "sgipc_convert_f64_to_f32": (ci, [vp, vp, vp, csz]),
    "sgipc_spmat_create": (ci, [vp, ctypes.POINTER(vp), vp, vp, vp, ci, ci, ci]),
    "sgipc_spmat_destroy": (ci, [vp]),
    "sgipc_spmat_apply": (ci, [vp, cd, vp, cd, vp]),
    "sgipc_spmat_apply_f32": (ci, [vp, cf, vp, cf, vp]),
    "sgipc_spmat_apply_f32_values": (ci, [vp, cd, vp, cd, vp]),
    "sgipc_spmat_layout": (ci, [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci)]),
    "sgipc_spmat_long_rows": (ci, [vp, ctypes.POINTER(ci)]),
    "sgipc_spmv_time": (ci, [vp, ci, vp, vp, vp, ci, ci, ci, ci, ctypes.POINTER(cd), ctypes.POINTER(ci)]),
    "sgipc_mas_time_apply": (ci, [vp, vp, vp, ci, ctypes.POINTER(cd)]),
    "sgipc_mas_create": (ci, [vp, ctypes.POINTER(vp), ci]),
    "sgipc_mas_destroy": (ci, [vp]),
    "sgipc_mas_init_neighbor": (ci, [vp, ci, ci, ci, vp, vp, vp]),
    "sgipc_mas_init_matrix": (ci, [vp]),
    "sgipc_mas_set_preconditioner_bcoo": (ci, [vp, vp, vp, vp, ci, ci, vp, ci, ci]),
    "sgipc_mas_preconditioning": (ci, [vp, vp, vp]),
    "sgipc_mas_levels": (ci, [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci)]),
    "sgipc_mas_tables": (ci, [vp, vp, vp]),
    "sgipc_mas_inverse_blocks": (ci, [vp, vp]),
    "sgipc_mas_level2_inverse": (ci, [vp, vp, ctypes.POINTER(ci)]),
    "sgipc_mas_partition_maps": (ci, [vp, ci, ci, ctypes.POINTER(ci), vp, vp]),
    "sgipc_partition_vertices": (ci, [ci, vp, vp, vp, ci, vp, ctypes.POINTER(ci)]),
    "sgipc_pcg_default_opts": (None, [ctypes.POINTER(PcgOpts)]),
    "sgipc_pcg_solve": (ci, [vp, vp, vp, vp, ci, ci, vp, vp, vp, ctypes.POINTER(PcgOpts), ctypes.POINTER(PcgStats)]),
    "sgipc_system_create": (ci, [vp, ctypes.POINTER(vp), ctypes.POINTER(SceneDesc), ci]),
    "sgipc_system_destroy": (ci, [vp]),
    "sgipc_system_solve": (ci, [vp, vp, vp, vp, ctypes.POINTER(PcgOpts), ctypes.POINTER(PcgStats),
                                ctypes.POINTER(SolveTiming)]),
    "sgipc_system_bcoo": (ci, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(ci)]),
    "sgipc_system_mas": (vp, [vp]),
    "sgipc_system_contacts": (ci, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(ci)]),
    "sgipc_system_energy": (ci, [vp, vp, vp, ctypes.POINTER(cd)]),
    "sgipc_system_line_search": (ci, [vp, vp, vp, cd, cd, vp, ctypes.POINTER(cd), ctypes.POINTER(cd),
                                      ctypes.POINTER(ci)]),
    "sgipc_mas_set_coarse_mode": (ci, [vp, ci]),
    "sgipc_mas_set_smoothing_weight": (ci, [vp, cd]),
    "sgipc_mas_set_level0_half": (ci, [vp, ci]),
    "sgipc_mpnn_create": (ci, [vp, ci, ci, ctypes.POINTER(vp)]),
    "sgipc_mpnn_destroy": (ci, [vp]),
    "sgipc_mpnn_num_params": (ci, [vp]),
    "sgipc_mpnn_set_params": (ci, [vp, vp, ci]),
    "sgipc_mpnn_get_params": (ci, [vp, vp, ci]),
    "sgipc_mpnn_init_params": (ci, [vp, ctypes.c_ulonglong]),
    "sgipc_mpnn_data_create": (ci, [vp, ci, vp, vp, vp, ci, ctypes.POINTER(vp)]),
    "sgipc_mpnn_data_add": (ci, [vp, vp, vp, ci]),
    "sgipc_mpnn_data_size": (ci, [vp]),
    "sgipc_mpnn_data_nodes": (ci, [vp]),
    "sgipc_mpnn_data_samples": (ci, [vp, ci, ci, vp, vp]),
    "sgipc_mpnn_data_graph": (ci, [vp, vp, vp, vp, vp]),
    "sgipc_mpnn_data_graph_nnz": (ci, [vp]),
    "sgipc_mpnn_data_destroy": (ci, [vp]),
    "sgipc_mpnn_forward": (ci, [vp, vp, ci, ci, vp]),
    "sgipc_mpnn_loss_grad": (ci, [vp, vp, ci, ci, ctypes.POINTER(cd), vp]),
    "sgipc_mpnn_train": (ci, [vp, vp, ci, ci, ctypes.c_float, ctypes.c_uint, vp]),
    "sgipc_system_collect": (ci, [vp, vp, ctypes.POINTER(PcgOpts), ci, ci, ctypes.POINTER(vp)]),
    "sgipc_system_set_learned": (ci, [vp, vp, ci]),
}

_LIB = None


def exported_symbols():
    """Names the binding expects the library to export (every ABI entry point)."""
    return list(_SIGS)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class SgipcError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def check(status, ctx=None, allow=()):
    if status == SGIPC_OK or status in allow:
        return status
    msg = lib().sgipc_last_error(ctx)
    raise SgipcError(status, msg.decode() if msg else "")


class DevArray:
    """Device allocation owned by a Context (freed with it or explicitly)."""

    def __init__(self, ctx, nbytes, dtype=None, shape=None):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype) if dtype is not None else None
        self.shape = shape
        p = vp()
        check(lib().sgipc_malloc(ctx.h, ctypes.byref(p), max(self.nbytes, 1)), ctx.h)
        self.ptr = p.value
        ctx._arrays.add(self)

    def free(self):
        if self.ptr and self.ctx.h:
            lib().sgipc_free(self.ctx.h, vp(self.ptr))
            self.ctx._arrays.discard(self)
        self.ptr = None

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(lib().sgipc_memcpy_dtoh(self.ctx.h, out.ctypes.data, vp(self.ptr), self.nbytes), self.ctx.h)
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        if self.nbytes:
            check(lib().sgipc_memcpy_htod(self.ctx.h, vp(self.ptr), a.ctypes.data, self.nbytes), self.ctx.h)
        return self

    def zero(self):
        if self.nbytes:
            check(lib().sgipc_memset(self.ctx.h, vp(self.ptr), 0, self.nbytes), self.ctx.h)
        return self

    @property
    def vp(self):
        return vp(self.ptr)


class Context:
    """One device context (HIP stream + workspaces) of the library."""

    def __init__(self, device=0):
        h = vp()
        check(lib().sgipc_create(ctypes.byref(h), int(device)))
        self.h = h
        self._arrays = set()
        self._deps = weakref.WeakSet()  # objects that use this context: closed first

    def adopt(self, obj):
        """Register an object (with a close() method) that uses this context;
        close() closes it before the context goes."""
        self._deps.add(obj)

    def empty(self, shape, dtype):
        shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        return DevArray(self, n, dtype, shape)

    def to_device(self, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = DevArray(self, a.nbytes, a.dtype, a.shape)
        d.upload(a)
        return d

    def synchronize(self):
        check(lib().sgipc_synchronize(self.h), self.h)

    def close(self):
        if self.h:
            for d in list(self._deps):
                d.close()
            for a in list(self._arrays):
                a.free()
            lib().sgipc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
