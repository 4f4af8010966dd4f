"""CPU ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over ``liboracle.so`` (built from ``sgipc_oracle.c`` by the
Makefile in this folder).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker; the product library never uses it.

Every function here restates the reference algorithm cited in the C source;
see ``sgipc_oracle.h`` for what pins each one.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

c_int_p = ctypes.POINTER(ctypes.c_int)
c_uint_p = ctypes.POINTER(ctypes.c_uint)
c_dbl_p = ctypes.POINTER(ctypes.c_double)


def _ptr(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        L.orc_tet_snh.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p,
                                  ctypes.c_double, ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_int_p]
        L.orc_tet_snh_energy.argtypes = [c_dbl_p, c_dbl_p, ctypes.c_double, ctypes.c_double]
        L.orc_tet_snh_energy.restype = ctypes.c_double
        L.orc_snh_dpdf.argtypes = [c_dbl_p, ctypes.c_double, ctypes.c_double, c_dbl_p, c_dbl_p]
        L.orc_barrier_pt.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, c_dbl_p, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_int_p]
        L.orc_barrier_energy.argtypes = [ctypes.c_double] * 3
        L.orc_barrier_energy.restype = ctypes.c_double
        L.orc_barrier_lambda0.argtypes = [ctypes.c_double] * 3
        L.orc_barrier_lambda0.restype = ctypes.c_double
        L.orc_mass_diag.argtypes = [ctypes.c_int, c_dbl_p, ctypes.c_int, c_dbl_p, c_int_p, c_int_p]
        L.orc_convert.argtypes = [ctypes.c_long, c_dbl_p, c_int_p, c_int_p, c_dbl_p, c_int_p, c_int_p]
        L.orc_convert.restype = ctypes.c_long
        L.orc_spmv.argtypes = [ctypes.c_double, ctypes.c_long, c_dbl_p, c_int_p, c_int_p, c_dbl_p,
                               ctypes.c_double, c_dbl_p, ctypes.c_int]
        L.orc_mas_create.argtypes = [ctypes.c_int, ctypes.c_int, c_uint_p, c_uint_p, c_uint_p, ctypes.c_int]
        L.orc_mas_create.restype = ctypes.c_void_p
        L.orc_mas_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_mas_build.argtypes = [ctypes.c_void_p, ctypes.c_long, c_dbl_p, c_int_p, c_int_p,
                                    ctypes.c_int, c_int_p, ctypes.c_int]
        L.orc_mas_apply.argtypes = [ctypes.c_void_p, c_dbl_p, c_dbl_p, ctypes.c_int]
        L.orc_mas_levels.argtypes = [ctypes.c_void_p, c_int_p, c_int_p, c_int_p]
        L.orc_mas_levels.restype = ctypes.c_int
        L.orc_mas_tables.argtypes = [ctypes.c_void_p, c_int_p, c_int_p]
        L.orc_mas_matrices.argtypes = [ctypes.c_void_p, c_dbl_p, ctypes.c_int]
        L.orc_mas_num_levels_for.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_mas_num_levels_for.restype = ctypes.c_int
        L.orc_pcg.argtypes = [ctypes.c_long, c_dbl_p, c_int_p, c_int_p, ctypes.c_int, ctypes.c_void_p,
                              c_dbl_p, c_dbl_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int, c_dbl_p]
        L.orc_pcg.restype = ctypes.c_int
        L.orc_set_fault.argtypes = [ctypes.c_int]
        L.orc_abd_body.argtypes = [ctypes.c_int, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p, ctypes.c_double,
                                   ctypes.c_double, c_int_p, ctypes.c_double, ctypes.c_double, c_dbl_p, c_dbl_p,
                                   c_int_p, c_int_p]
        L.orc_abd_shape_energy.argtypes = [c_dbl_p]
        L.orc_abd_shape_energy.restype = ctypes.c_double
        L.orc_abd_shape_grad_hess.argtypes = [c_dbl_p, c_dbl_p, c_dbl_p]
        L.orc_abd_contact_map.argtypes = [ctypes.c_long, c_dbl_p, c_int_p, c_int_p, ctypes.c_int, c_int_p, c_dbl_p,
                                          ctypes.c_int, c_int_p, c_int_p, c_dbl_p, c_int_p, c_int_p]
        L.orc_dirichlet_zero.argtypes = [ctypes.c_long, c_int_p, c_int_p, ctypes.c_int, c_int_p, c_dbl_p]
        L.orc_abd_contact_map.restype = ctypes.c_long
        L.orc_abd_jacobi_build.argtypes = [ctypes.c_long, c_dbl_p, c_int_p, c_int_p, ctypes.c_int, c_dbl_p]
        L.orc_abd_jacobi_build.restype = ctypes.c_int
        L.orc_pcg2.argtypes = [ctypes.c_long, c_dbl_p, c_int_p, c_int_p, ctypes.c_int, ctypes.c_void_p, c_dbl_p,
                               c_dbl_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, c_dbl_p]
        L.orc_pcg2.restype = ctypes.c_int
        L.orc_tri_bw.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, c_dbl_p, c_dbl_p] + [ctypes.c_double] * 4 + \
            [ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_int_p]
        L.orc_tri_bw_energy.argtypes = [c_dbl_p, c_dbl_p] + [ctypes.c_double] * 4
        L.orc_tri_bw_energy.restype = ctypes.c_double
        L.orc_bw_dpdf.argtypes = [c_dbl_p] + [ctypes.c_double] * 3 + [c_dbl_p, c_dbl_p]
        L.orc_quad_bending_q.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, c_int_p, c_dbl_p]
        L.orc_quad_bending.argtypes = [ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_int_p, c_dbl_p,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, c_dbl_p, c_dbl_p,
                                       c_int_p, c_int_p]
        L.orc_contact_detect.argtypes = [ctypes.c_int, c_dbl_p, c_dbl_p, ctypes.c_int, c_int_p, ctypes.c_int,
                                         c_int_p, ctypes.c_int, c_int_p, c_int_p, c_int_p, ctypes.c_double,
                                         ctypes.c_int, c_int_p, c_int_p, c_int_p]
        L.orc_contact_detect.restype = ctypes.c_int
        L.orc_barrier.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, c_int_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_int_p]
        L.orc_barrier.restype = ctypes.c_int
        L.orc_barrier_energy_pairs.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, ctypes.c_double, ctypes.c_double]
        L.orc_barrier_energy_pairs.restype = ctypes.c_double
        L.orc_friction_prepare.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, ctypes.c_double, ctypes.c_double, c_dbl_p,
                                           c_dbl_p, c_dbl_p]
        L.orc_friction.argtypes = [ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_int_p, ctypes.c_int, ctypes.c_int, c_dbl_p,
                                   c_dbl_p, c_dbl_p, ctypes.c_double, ctypes.c_double, ctypes.c_int, c_dbl_p, c_dbl_p,
                                   c_int_p, c_int_p]
        L.orc_friction_energy.argtypes = [ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_dbl_p, c_dbl_p, c_dbl_p,
                                          ctypes.c_double, ctypes.c_double]
        L.orc_friction_energy.restype = ctypes.c_double
        L.orc_ground_friction_prepare.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, c_dbl_p, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_double, c_dbl_p]
        L.orc_ground_friction.argtypes = [ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_dbl_p, c_dbl_p, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_int_p]
        L.orc_ground_friction_energy.argtypes = [ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_dbl_p, c_dbl_p,
                                                 ctypes.c_double, ctypes.c_double]
        L.orc_ground_friction_energy.restype = ctypes.c_double
        L.orc_ccd_self_step.argtypes = [ctypes.c_int, c_dbl_p, c_dbl_p, ctypes.c_double, ctypes.c_int, c_int_p,
                                        ctypes.c_int, c_int_p, ctypes.c_int, c_int_p, c_int_p, c_int_p,
                                        ctypes.c_double, ctypes.c_double, c_int_p]
        L.orc_ccd_self_step.restype = ctypes.c_double
        L.orc_ground_step.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, c_dbl_p, ctypes.c_double, c_dbl_p,
                                      ctypes.c_double]
        L.orc_ground_step.restype = ctypes.c_double
        L.orc_injective_step.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, c_dbl_p, ctypes.c_double, ctypes.c_double]
        L.orc_injective_step.restype = ctypes.c_double
        L.orc_energy_tets.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, c_dbl_p, c_dbl_p, c_dbl_p, c_dbl_p,
                                      ctypes.c_double]
        L.orc_energy_tets.restype = ctypes.c_double
        L.orc_energy_tris.argtypes = [ctypes.c_int, c_dbl_p, c_int_p, c_dbl_p, c_dbl_p] + [ctypes.c_double] * 4
        L.orc_energy_tris.restype = ctypes.c_double
        L.orc_energy_bending.argtypes = [ctypes.c_int, c_dbl_p, c_dbl_p, c_int_p, c_int_p, c_dbl_p, ctypes.c_double,
                                         ctypes.c_double]
        L.orc_energy_bending.restype = ctypes.c_double
        _LIB = L
    return _LIB


def set_threads(k: int):
    """OpenMP threads of the oracle's parallel loops."""
    lib().orc_set_threads(int(k))


def set_fault(k: int):
    lib().orc_set_fault(int(k))


# ---------------------------------------------------------------- element terms
def tet_snh(scene, hess_offset=0, with_grad=True):
    L = lib()
    m = scene.n_tets
    vals = np.zeros((10 * m, 9))
    rows = np.zeros(10 * m, dtype=np.int32)
    cols = np.zeros(10 * m, dtype=np.int32)
    grad = np.zeros(3 * scene.n) if with_grad else None
    x = np.ascontiguousarray(scene.x, dtype=np.float64)
    tets = np.ascontiguousarray(scene.tets, dtype=np.int32)
    L.orc_tet_snh(m, _ptr(x, c_dbl_p), _ptr(tets, c_int_p), _ptr(scene.dm_inv, c_dbl_p),
                  _ptr(scene.vol, c_dbl_p), _ptr(scene.mu, c_dbl_p), _ptr(scene.lam, c_dbl_p),
                  scene.dt, hess_offset, _ptr(grad, c_dbl_p), _ptr(vals, c_dbl_p),
                  _ptr(rows, c_int_p), _ptr(cols, c_int_p))
    return vals, rows, cols, grad


def barrier_pt(scene, hess_offset=0, with_grad=True):
    L = lib()
    p = scene.n_pairs
    vals = np.zeros((16 * p, 9))
    rows = np.zeros(16 * p, dtype=np.int32)
    cols = np.zeros(16 * p, dtype=np.int32)
    grad = np.zeros(3 * scene.n) if with_grad else None
    if p:
        pairs = np.ascontiguousarray(scene.pairs, dtype=np.int32)
        px = None if scene.pair_x is None else np.ascontiguousarray(scene.pair_x, dtype=np.float64)
        L.orc_barrier_pt(p, _ptr(np.ascontiguousarray(scene.x), c_dbl_p), _ptr(pairs, c_int_p),
                         _ptr(px, c_dbl_p), scene.dhat ** 2, scene.kappa, hess_offset,
                         _ptr(grad, c_dbl_p), _ptr(vals, c_dbl_p), _ptr(rows, c_int_p), _ptr(cols, c_int_p))
    return vals, rows, cols, grad


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def contact_detect(x, rest, faces, edges, surf, dhat_sq, body_id=None, btype=None, capacity=None):
    """Contact pairs by brute force (contact_oracle.c): (pairs (m, 4) int32,
    mat_index (m,), counts [total, 0, PP, PE, PT+EE])."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    rest = np.ascontiguousarray(rest, dtype=np.float64).reshape(-1)
    faces, edges, surf = _i32(faces).reshape(-1, 3), _i32(edges).reshape(-1, 2), _i32(surf).reshape(-1)
    body_id, btype = _i32(body_id), _i32(btype)
    cap = capacity if capacity is not None else 1 << 16
    while True:
        pairs = np.zeros((cap, 4), dtype=np.int32)
        mi = np.zeros(cap, dtype=np.int32)
        counts = np.zeros(5, dtype=np.int32)
        st = lib().orc_contact_detect(x.size // 3, _ptr(x, c_dbl_p), _ptr(rest, c_dbl_p), faces.shape[0],
                                      _ptr(faces, c_int_p), edges.shape[0], _ptr(edges, c_int_p), surf.size,
                                      _ptr(surf, c_int_p), _ptr(body_id, c_int_p), _ptr(btype, c_int_p),
                                      float(dhat_sq), cap, _ptr(pairs, c_int_p), _ptr(mi, c_int_p),
                                      _ptr(counts, c_int_p))
        if st == 0:
            n = int(counts[0])
            return pairs[:n].copy(), mi[:n].copy(), counts
        if capacity is not None:
            return None, None, counts
        cap = int(counts[0])


def barrier(x, pairs, mat_index, counts, dhat_sq, kappa, hess_offset=0, with_grad=True):
    """Barrier gradient and Hessian of every pair type in the contact arena
    layout [PT+EE x 16 | PE x 9 | PP x 4]."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    pairs = _i32(pairs).reshape(-1, 4)
    mat_index = _i32(mat_index)
    n4, n3, n2 = int(counts[4]), int(counts[3]), int(counts[2])
    nt = 16 * n4 + 9 * n3 + 4 * n2
    vals = np.zeros((nt, 9))
    rows = np.zeros(nt, dtype=np.int32)
    cols = np.zeros(nt, dtype=np.int32)
    grad = np.zeros(x.size) if with_grad else None
    st = lib().orc_barrier(pairs.shape[0], _ptr(x, c_dbl_p), _ptr(pairs, c_int_p), _ptr(mat_index, c_int_p), n4, n3,
                           float(dhat_sq), float(kappa), hess_offset, _ptr(grad, c_dbl_p), _ptr(vals, c_dbl_p),
                           _ptr(rows, c_int_p), _ptr(cols, c_int_p))
    if st != 0:
        raise ValueError("smooth-encoded pair: not produced by the detection")
    return vals, rows, cols, grad


def barrier_energy(x, pairs, dhat_sq, kappa):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    pairs = _i32(pairs).reshape(-1, 4)
    return lib().orc_barrier_energy_pairs(pairs.shape[0], _ptr(x, c_dbl_p), _ptr(pairs, c_int_p), float(dhat_sq),
                                          float(kappa))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


def friction_prepare(xt, pairs, dhat_sq, kappa):
    """(lambda, coord (n, 2), basis (n, 6)) of the pairs at the step start xt."""
    xt, pairs = _f64(xt), _i32(pairs).reshape(-1, 4)
    n = pairs.shape[0]
    lam, coord, basis = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 6))
    if lib().orc_friction_prepare(n, _ptr(xt, c_dbl_p), _ptr(pairs, c_int_p), float(dhat_sq), float(kappa),
                                  _ptr(lam, c_dbl_p), _ptr(coord, c_dbl_p), _ptr(basis, c_dbl_p)) != 0:
        raise ValueError("smooth-encoded pair")
    return lam, coord, basis


def friction(x, xt, pairs, mat_index, counts, lam, coord, basis, eps2, mu, hess_offset=0, with_grad=True):
    x, xt, pairs = _f64(x), _f64(xt), _i32(pairs).reshape(-1, 4)
    n4, n3, n2 = int(counts[4]), int(counts[3]), int(counts[2])
    nt = 16 * n4 + 9 * n3 + 4 * n2
    vals, rows, cols = np.zeros((nt, 9)), np.zeros(nt, np.int32), np.zeros(nt, np.int32)
    grad = np.zeros(x.size) if with_grad else None
    lam, coord, basis = _f64(lam), _f64(coord), _f64(basis)
    lib().orc_friction(pairs.shape[0], _ptr(x, c_dbl_p), _ptr(xt, c_dbl_p), _ptr(pairs, c_int_p),
                       _ptr(_i32(mat_index), c_int_p), n4, n3, _ptr(lam, c_dbl_p), _ptr(coord, c_dbl_p),
                       _ptr(basis, c_dbl_p), float(eps2), float(mu), hess_offset, _ptr(grad, c_dbl_p),
                       _ptr(vals, c_dbl_p), _ptr(rows, c_int_p), _ptr(cols, c_int_p))
    return vals, rows, cols, grad


def friction_energy(x, xt, pairs, lam, coord, basis, eps2, mu):
    x, xt, pairs = _f64(x), _f64(xt), _i32(pairs).reshape(-1, 4)
    lam, coord, basis = _f64(lam), _f64(coord), _f64(basis)
    return lib().orc_friction_energy(pairs.shape[0], _ptr(x, c_dbl_p), _ptr(xt, c_dbl_p), _ptr(pairs, c_int_p),
                                     _ptr(lam, c_dbl_p), _ptr(coord, c_dbl_p), _ptr(basis, c_dbl_p), float(eps2),
                                     float(mu))


def ground_friction_prepare(xt, verts, normal, offset, dhat_sq, kappa):
    xt, verts, normal = _f64(xt), _i32(verts), _f64(normal)
    lam = np.zeros(verts.size)
    lib().orc_ground_friction_prepare(verts.size, _ptr(xt, c_dbl_p), _ptr(verts, c_int_p), _ptr(normal, c_dbl_p),
                                      float(offset), float(dhat_sq), float(kappa), _ptr(lam, c_dbl_p))
    return lam


def ground_friction(x, xt, verts, normal, lam, eps2, mu, hess_offset=0, with_grad=True):
    x, xt, verts, normal, lam = _f64(x), _f64(xt), _i32(verts), _f64(normal), _f64(lam)
    n = verts.size
    vals, rows, cols = np.zeros((n, 9)), np.zeros(n, np.int32), np.zeros(n, np.int32)
    grad = np.zeros(x.size) if with_grad else None
    lib().orc_ground_friction(n, _ptr(x, c_dbl_p), _ptr(xt, c_dbl_p), _ptr(verts, c_int_p), _ptr(normal, c_dbl_p),
                              _ptr(lam, c_dbl_p), float(eps2), float(mu), hess_offset, _ptr(grad, c_dbl_p),
                              _ptr(vals, c_dbl_p), _ptr(rows, c_int_p), _ptr(cols, c_int_p))
    return vals, rows, cols, grad


def ground_friction_energy(x, xt, verts, normal, lam, eps2, mu):
    x, xt, verts, normal, lam = _f64(x), _f64(xt), _i32(verts), _f64(normal), _f64(lam)
    return lib().orc_ground_friction_energy(verts.size, _ptr(x, c_dbl_p), _ptr(xt, c_dbl_p), _ptr(verts, c_int_p),
                                            _ptr(normal, c_dbl_p), _ptr(lam, c_dbl_p), float(eps2), float(mu))


def ccd_self_step(x, direction, alpha, faces, edges, surf, dhat_sq, slackness, body_id=None, btype=None):
    """Largest feasible fraction of the step x - t direction (additive CCD
    over the swept-box candidates); returns (step, candidate count)."""
    x, d = _f64(x), _f64(direction)
    faces, edges, surf = _i32(faces).reshape(-1, 3), _i32(edges).reshape(-1, 2), _i32(surf)
    nc = ctypes.c_int(0)
    st = lib().orc_ccd_self_step(x.size // 3, _ptr(x, c_dbl_p), _ptr(d, c_dbl_p), float(alpha), faces.shape[0],
                                 _ptr(faces, c_int_p), edges.shape[0], _ptr(edges, c_int_p), surf.size,
                                 _ptr(surf, c_int_p), _ptr(_i32(body_id), c_int_p), _ptr(_i32(btype), c_int_p),
                                 float(dhat_sq), float(slackness), ctypes.byref(nc))
    return st, nc.value


def ground_step(x, surf, normal, offset, direction, slackness):
    x, surf, normal, d = _f64(x), _i32(surf), _f64(normal), _f64(direction)
    return lib().orc_ground_step(surf.size, _ptr(x, c_dbl_p), _ptr(surf, c_int_p), _ptr(normal, c_dbl_p),
                                 float(offset), _ptr(d, c_dbl_p), float(slackness))


def injective_step(x, tets, direction, slackness, error_rate=1e-20):
    x, tets, d = _f64(x), _i32(tets).reshape(-1, 4), _f64(direction)
    return lib().orc_injective_step(tets.shape[0], _ptr(x, c_dbl_p), _ptr(tets, c_int_p), _ptr(d, c_dbl_p),
                                    float(slackness), float(error_rate))


def scene_energy(scene, x, x_tilde=None):
    """Newton objective at positions x (n, 3): inertia (with x_tilde) +
    elastic terms + the barrier of the pairs at x (detected at x for scenes
    with a contact surface, else the seeded pairs)."""
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    xf = x.reshape(-1)
    e = 0.0
    if x_tilde is not None:
        dx = x - np.asarray(x_tilde).reshape(-1, 3)
        e += 0.5 * float(np.sum(scene.mass * np.sum(dx * dx, axis=1)))
    if getattr(scene, "n_tets", 0):
        e += L.orc_energy_tets(scene.n_tets, _ptr(xf, c_dbl_p), _ptr(_i32(scene.tets).reshape(-1), c_int_p),
                               _ptr(_f64(scene.dm_inv), c_dbl_p), _ptr(_f64(scene.vol), c_dbl_p),
                               _ptr(_f64(scene.mu), c_dbl_p), _ptr(_f64(scene.lam), c_dbl_p), scene.dt)
    if getattr(scene, "n_tris", 0):
        e += L.orc_energy_tris(scene.n_tris, _ptr(xf, c_dbl_p), _ptr(_i32(scene.tris).reshape(-1), c_int_p),
                               _ptr(_f64(scene.dm_inv2), c_dbl_p), _ptr(_f64(scene.area), c_dbl_p), scene.k_stretch,
                               scene.k_shear, scene.strain_rate, scene.dt)
    if getattr(scene, "n_hinges", 0):
        q = np.zeros((scene.n_hinges, 16))
        edges, adj = _i32(scene.edges).reshape(-1), _i32(scene.adj).reshape(-1)
        rest = _f64(scene.rest)
        L.orc_quad_bending_q(scene.n_hinges, _ptr(rest, c_dbl_p), _ptr(edges, c_int_p), _ptr(adj, c_int_p),
                             _ptr(q, c_dbl_p))
        e += L.orc_energy_bending(scene.n_hinges, _ptr(xf, c_dbl_p), _ptr(rest, c_dbl_p), _ptr(edges, c_int_p),
                                  _ptr(adj, c_int_p), _ptr(q, c_dbl_p), scene.k_bend, scene.dt)
    dsq = scene.dhat ** 2
    if getattr(scene, "contact_surface", None) is not None:
        cs = scene.contact_surface
        pr, _, _ = contact_detect(x, scene.rest, cs.faces, cs.edges, cs.surf, dsq, body_id=cs.body_id)
        e += barrier_energy(x, pr, dsq, scene.kappa)
    elif scene.n_pairs:
        if scene.pair_x is not None:
            px = np.ascontiguousarray(scene.pair_x, dtype=np.float64).reshape(-1, 3)
            loc = np.tile(np.array([-1, 1, 2, 3], np.int32), (scene.n_pairs, 1))
            loc[:, 1:] += 4 * np.arange(scene.n_pairs)[:, None]
            loc[:, 0] = -(4 * np.arange(scene.n_pairs)) - 1
            e += barrier_energy(px, loc, dsq, scene.kappa)
        else:
            e += barrier_energy(x, scene.pairs, dsq, scene.kappa)
    return e


def line_search(scene, direction, x_tilde=None, alpha=1.0, slackness=0.8):
    """The system's line search restated: feasible step (CCD for scenes with
    a contact surface, tet injectivity), then halving while the energy
    exceeds the start's, at most 8 times.  Returns (alpha, x, E0, E, stopped)."""
    x0 = np.ascontiguousarray(scene.x, dtype=np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(direction, dtype=np.float64).reshape(-1, 3)
    a = alpha
    cs = getattr(scene, "contact_surface", None)
    if cs is not None:
        a = min(a, ccd_self_step(x0, d, a, cs.faces, cs.edges, cs.surf, scene.dhat ** 2, slackness,
                                 body_id=cs.body_id)[0])
    if getattr(scene, "n_tets", 0):
        a = min(a, injective_step(x0, scene.tets, d, slackness))
    e0 = scene_energy(scene, x0, x_tilde)
    e = scene_energy(scene, x0 - a * d, x_tilde)
    k = 0
    while e > e0 and k <= 8:
        a *= 0.5
        k += 1
        e = scene_energy(scene, x0 - a * d, x_tilde)
    if k > 8:
        return a, x0.copy(), e0, e, True
    return a, x0 - a * d, e0, e, False


def tri_bw(scene, hess_offset=0, with_grad=True):
    """Baraff-Witkin cloth triangles: 6 blocks per triangle."""
    L = lib()
    m = scene.n_tris
    vals = np.zeros((6 * m, 9))
    rows = np.zeros(6 * m, dtype=np.int32)
    cols = np.zeros(6 * m, dtype=np.int32)
    grad = np.zeros(3 * scene.n) if with_grad else None
    if m:
        L.orc_tri_bw(m, _ptr(np.ascontiguousarray(scene.x), c_dbl_p),
                     _ptr(np.ascontiguousarray(scene.tris, dtype=np.int32), c_int_p),
                     _ptr(np.ascontiguousarray(scene.dm_inv2), c_dbl_p), _ptr(np.ascontiguousarray(scene.area), c_dbl_p),
                     scene.k_stretch, scene.k_shear, scene.strain_rate, scene.dt, hess_offset,
                     _ptr(grad, c_dbl_p), _ptr(vals, c_dbl_p), _ptr(rows, c_int_p), _ptr(cols, c_int_p))
    return vals, rows, cols, grad


def tri_bw_energy(x9, dm_inv2, area, k_stretch, k_shear, rate):
    x9 = np.ascontiguousarray(x9, dtype=np.float64)
    d = np.ascontiguousarray(dm_inv2, dtype=np.float64)
    return lib().orc_tri_bw_energy(_ptr(x9, c_dbl_p), _ptr(d, c_dbl_p), area, k_stretch, k_shear, rate)


def bw_dpdf(F6, k_stretch, k_shear, rate):
    F6 = np.ascontiguousarray(F6, dtype=np.float64)
    P, H = np.zeros(6), np.zeros(36)
    lib().orc_bw_dpdf(_ptr(F6, c_dbl_p), k_stretch, k_shear, rate, _ptr(P, c_dbl_p), _ptr(H, c_dbl_p))
    return P, H.reshape(6, 6)


def quad_bending_q(rest_x, edges, adj):
    e = np.ascontiguousarray(edges, dtype=np.int32)
    a = np.ascontiguousarray(adj, dtype=np.int32)
    Q = np.zeros((e.shape[0], 16))
    lib().orc_quad_bending_q(e.shape[0], _ptr(np.ascontiguousarray(rest_x, dtype=np.float64), c_dbl_p),
                             _ptr(e, c_int_p), _ptr(a, c_int_p), _ptr(Q, c_dbl_p))
    return Q


def quad_bending(scene, Q=None, hess_offset=0, with_grad=True):
    """Quadratic bending hinges: 10 blocks per hinge."""
    L = lib()
    m = scene.n_hinges
    vals = np.zeros((10 * m, 9))
    rows = np.zeros(10 * m, dtype=np.int32)
    cols = np.zeros(10 * m, dtype=np.int32)
    grad = np.zeros(3 * scene.n) if with_grad else None
    if m:
        if Q is None:
            Q = quad_bending_q(scene.rest, scene.edges, scene.adj)
        L.orc_quad_bending(m, _ptr(np.ascontiguousarray(scene.x), c_dbl_p),
                           _ptr(np.ascontiguousarray(scene.rest), c_dbl_p),
                           _ptr(np.ascontiguousarray(scene.edges, dtype=np.int32), c_int_p),
                           _ptr(np.ascontiguousarray(scene.adj, dtype=np.int32), c_int_p),
                           _ptr(np.ascontiguousarray(Q), c_dbl_p), scene.k_bend, scene.dt, hess_offset,
                           _ptr(grad, c_dbl_p), _ptr(vals, c_dbl_p), _ptr(rows, c_int_p), _ptr(cols, c_int_p))
    return vals, rows, cols, grad


def abd_body(scene, with_grad=True):
    """Affine-body terms: 10 blocks per body at rows 4b..4b+3."""
    nb = scene.n_bodies
    vals = np.zeros((10 * nb, 9))
    rows = np.zeros(10 * nb, dtype=np.int32)
    cols = np.zeros(10 * nb, dtype=np.int32)
    grad = np.zeros(12 * nb) if with_grad else None
    if nb:
        fixed = np.ascontiguousarray(scene.body_fixed, dtype=np.int32)
        lib().orc_abd_body(nb, _ptr(np.ascontiguousarray(scene.q), c_dbl_p),
                           _ptr(np.ascontiguousarray(scene.q_tilde), c_dbl_p),
                           _ptr(np.ascontiguousarray(scene.mass10), c_dbl_p),
                           _ptr(np.ascontiguousarray(scene.body_volume), c_dbl_p), scene.kappa_abd, scene.dt,
                           _ptr(fixed, c_int_p), getattr(scene, "motor_speed", 31.4),
                           getattr(scene, "motor_strength", 10.0), _ptr(grad, c_dbl_p), _ptr(vals, c_dbl_p), _ptr(rows, c_int_p),
                           _ptr(cols, c_int_p))
    return vals, rows, cols, grad


def abd_shape(q12):
    q12 = np.ascontiguousarray(q12, dtype=np.float64)
    g, H = np.zeros(9), np.zeros(81)
    e = lib().orc_abd_shape_energy(_ptr(q12, c_dbl_p))
    lib().orc_abd_shape_grad_hess(_ptr(q12, c_dbl_p), _ptr(g, c_dbl_p), _ptr(H, c_dbl_p))
    return e, g, H.reshape(9, 9)


def abd_contact_map(scene, vals, rows, cols):
    """Vertex-level contact triplets -> system triplets by class (FEM-FEM,
    ABD-FEM, ABD-ABD)."""
    n = rows.size
    cap = 16 * max(n, 1)
    ov = np.zeros((cap, 9))
    orr = np.zeros(cap, dtype=np.int32)
    oc = np.zeros(cap, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    m = lib().orc_abd_contact_map(n, _ptr(vals, c_dbl_p), _ptr(rows, c_int_p), _ptr(cols, c_int_p),
                                  scene.n_abd_points, _ptr(np.ascontiguousarray(scene.point_body, dtype=np.int32), c_int_p),
                                  _ptr(np.ascontiguousarray(scene.x_bar), c_dbl_p), scene.n_bodies,
                                  _ptr(np.ascontiguousarray(scene.body_fixed, dtype=np.int32), c_int_p),
                                  _ptr(_btype(scene), c_int_p), _ptr(ov, c_dbl_p), _ptr(orr, c_int_p),
                                  _ptr(oc, c_int_p))
    return ov[:m].copy(), orr[:m].copy(), oc[:m].copy()


def abd_jacobi(vals, rows, cols, n_bodies):
    """Inverse 12x12 diagonal blocks (n_bodies, 12, 12)."""
    out = np.zeros((max(n_bodies, 1), 144))
    st = lib().orc_abd_jacobi_build(rows.size, _ptr(np.ascontiguousarray(vals, dtype=np.float64), c_dbl_p),
                                    _ptr(np.ascontiguousarray(rows, dtype=np.int32), c_int_p),
                                    _ptr(np.ascontiguousarray(cols, dtype=np.int32), c_int_p), n_bodies,
                                    _ptr(out, c_dbl_p))
    if st != 0:
        raise RuntimeError("singular body block")
    return out[:n_bodies].reshape(-1, 12, 12)


def mass_diag(scene, hess_offset=0):
    n = scene.n
    vals = np.zeros((n, 9))
    rows = np.zeros(n, dtype=np.int32)
    cols = np.zeros(n, dtype=np.int32)
    lib().orc_mass_diag(n, _ptr(np.ascontiguousarray(scene.mass), c_dbl_p), hess_offset,
                        _ptr(vals, c_dbl_p), _ptr(rows, c_int_p), _ptr(cols, c_int_p))
    return vals, rows, cols


def _btype(scene):
    bt = getattr(scene, "boundary_type", None)
    return None if bt is None else np.ascontiguousarray(bt, dtype=np.int32)


def dirichlet_zero(vals, rows, cols, row0, btype):
    """Zero the triplets with an end on a Dirichlet FEM vertex (in place)."""
    if btype is None or rows.size == 0:
        return vals
    lib().orc_dirichlet_zero(rows.size, _ptr(np.ascontiguousarray(rows, dtype=np.int32), c_int_p),
                             _ptr(np.ascontiguousarray(cols, dtype=np.int32), c_int_p), row0,
                             _ptr(btype, c_int_p), _ptr(vals, c_dbl_p))
    return vals


def point_gradient_to_system(scene, gp):
    """Contact gradient of the points (3 per point: body points, then FEM
    points) in the system rows: J^T g into the 12 rows of each non-fixed body
    (setup_abd_system_gradient_and_hessian.cu:349-385), g into the FEM vertex
    rows after the 4 n_bodies body rows."""
    nb, na = scene.n_bodies, scene.n_abd_points
    g = np.zeros(3 * scene.n_sys)
    g[12 * nb:] = gp[3 * na:]
    gpp = gp[:3 * na].reshape(na, 3)
    keep = np.asarray(scene.body_fixed)[scene.point_body] != 1
    for i in np.nonzero(keep)[0]:
        b = scene.point_body[i]
        jt = np.zeros(12)
        jt[:3] = gpp[i]
        jt[3:] = np.outer(gpp[i], scene.x_bar[i]).reshape(-1)   # row r of A: g_r xbar
        g[12 * b:12 * b + 12] += jt
    return g


def scene_pairs(scene, with_index=False):
    """The scene's contact pairs: detected from its contact surface at the
    current positions (contact_detect), else its seeded pairs."""
    cs = getattr(scene, "contact_surface", None)
    if cs is None:
        return (scene.pairs, None, None) if with_index else scene.pairs
    pr, mi, counts = contact_detect(scene.x, scene.rest, cs.faces, cs.edges, cs.surf, scene.dhat ** 2,
                                    body_id=cs.body_id, btype=getattr(scene, "boundary_type", None))
    return (pr, mi, counts) if with_index else pr


def triplets(scene, with_grad=False):
    """All triplets in the reference's arena order (GIPC.cu:10291-10485):
    contacts, affine bodies, tets, bending hinges, cloth triangles, mass.  On
    a scene with affine bodies the contact blocks are mapped to the system
    rows by class first, and the FEM rows start at 4 * n_bodies.  Triplets
    with an end on a Dirichlet vertex (scene.boundary_type) are zeroed, the
    mass diagonal excepted (GIPC.cu:10332-10363,10444-10462).  The gradient
    is the system right-hand side the subsystems assemble: body rows
    M(q - q~) + shape + J^T (contact gradient), FEM rows element + contact
    gradients, zero at Dirichlet vertices (fem_linear_subsystem.cu:56-84,
    abd_linear_subsystem.cu:20-32)."""
    nb = getattr(scene, "n_bodies", 0)
    off = 4 * nb
    bt = _btype(scene)
    if nb:
        bv, br, bc, _ = barrier_pt(scene, with_grad=False)
        gp = None
        if with_grad and scene.n_pairs:
            gp = np.zeros(3 * (scene.n_abd_points + scene.n))
            lib().orc_barrier_pt(scene.n_pairs, _ptr(np.ascontiguousarray(scene.x), c_dbl_p),
                                 _ptr(np.ascontiguousarray(scene.pairs, dtype=np.int32), c_int_p),
                                 _ptr(np.ascontiguousarray(scene.pair_x, dtype=np.float64), c_dbl_p),
                                 scene.dhat ** 2, scene.kappa, 0, _ptr(gp, c_dbl_p),
                                 _ptr(np.zeros((16 * scene.n_pairs, 9)), c_dbl_p),
                                 _ptr(np.zeros(16 * scene.n_pairs, dtype=np.int32), c_int_p),
                                 _ptr(np.zeros(16 * scene.n_pairs, dtype=np.int32), c_int_p))
        bv, br, bc = abd_contact_map(scene, bv, br, bc)
        bg = None
    elif getattr(scene, "contact_surface", None) is not None:
        pr, mi, counts = scene_pairs(scene, with_index=True)
        bv, br, bc, bg = barrier(scene.x, pr, mi, counts, scene.dhat ** 2, scene.kappa, with_grad=with_grad)
    else:
        bv, br, bc, bg = barrier_pt(scene, with_grad=with_grad)
    parts = [(bv, br, bc, bg)]
    if nb:
        parts.append(abd_body(scene, with_grad=with_grad))
    if getattr(scene, "n_tets", 0):
        parts.append(tet_snh(scene, hess_offset=off, with_grad=with_grad))
    if getattr(scene, "n_hinges", 0):
        parts.append(quad_bending(scene, hess_offset=off, with_grad=with_grad))
    if getattr(scene, "n_tris", 0):
        parts.append(tri_bw(scene, hess_offset=off, with_grad=with_grad))
    vals = np.concatenate([p[0] for p in parts])
    rows = np.concatenate([p[1] for p in parts])
    cols = np.concatenate([p[2] for p in parts])
    dirichlet_zero(vals, rows, cols, off, bt)
    mv, mr, mc = mass_diag(scene, hess_offset=off)
    vals = np.concatenate([vals, mv])
    rows = np.concatenate([rows, mr])
    cols = np.concatenate([cols, mc])
    grad = None
    if with_grad:
        if nb:
            grad = np.zeros(3 * scene.n_sys)
            grad[:12 * nb] = parts[1][3]
            for p in parts[2:]:
                grad[12 * nb:] += p[3]
            if gp is not None:
                grad += point_gradient_to_system(scene, gp)
        else:
            grad = sum(p[3] for p in parts)
        if bt is not None:
            grad[3 * off:].reshape(-1, 3)[bt != 0] = 0.0
    return vals, rows, cols, grad


def convert(vals, rows, cols):
    n = rows.size
    ov = np.zeros((max(n, 1), 9))
    orr = np.zeros(max(n, 1), dtype=np.int32)
    oc = np.zeros(max(n, 1), dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    nnz = lib().orc_convert(n, _ptr(vals, c_dbl_p), _ptr(rows, c_int_p), _ptr(cols, c_int_p),
                            _ptr(ov, c_dbl_p), _ptr(orr, c_int_p), _ptr(oc, c_int_p))
    return ov[:nnz].copy(), orr[:nnz].copy(), oc[:nnz].copy()


def assemble(scene):
    """Unique upper BCOO (vals (nnz,9) column-major blocks, rows, cols)."""
    v, r, c, _ = triplets(scene)
    return convert(v, r, c)


def assemble_arena(scene):
    """``assemble`` for tet scenes with seeded pairs, no bodies and no
    Dirichlet vertices, writing every element's triplets straight into one
    preallocated arena in the same order (contacts, tets, mass) instead of
    concatenating per-kind arrays: the CPU baseline's assembly."""
    assert getattr(scene, "n_bodies", 0) == 0 and getattr(scene, "contact_surface", None) is None
    assert _btype(scene) is None and getattr(scene, "n_tris", 0) == 0 and getattr(scene, "n_hinges", 0) == 0
    L = lib()
    p, m, n = scene.n_pairs, scene.n_tets, scene.n
    tot = 16 * p + 10 * m + n
    vals = np.empty((tot, 9))
    rows = np.empty(tot, dtype=np.int32)
    cols = np.empty(tot, dtype=np.int32)
    x = np.ascontiguousarray(scene.x, dtype=np.float64)

    def at(a, k, t):
        return ctypes.cast(a.ctypes.data + k * a.strides[0], t)

    if p:
        pairs = np.ascontiguousarray(scene.pairs, dtype=np.int32)
        px = None if scene.pair_x is None else np.ascontiguousarray(scene.pair_x, dtype=np.float64)
        L.orc_barrier_pt(p, _ptr(x, c_dbl_p), _ptr(pairs, c_int_p), _ptr(px, c_dbl_p), scene.dhat ** 2, scene.kappa,
                         0, None, at(vals, 0, c_dbl_p), at(rows, 0, c_int_p), at(cols, 0, c_int_p))
    tets = np.ascontiguousarray(scene.tets, dtype=np.int32)
    k = 16 * p
    L.orc_tet_snh(m, _ptr(x, c_dbl_p), _ptr(tets, c_int_p), _ptr(scene.dm_inv, c_dbl_p), _ptr(scene.vol, c_dbl_p),
                  _ptr(scene.mu, c_dbl_p), _ptr(scene.lam, c_dbl_p), scene.dt, 0, None, at(vals, k, c_dbl_p),
                  at(rows, k, c_int_p), at(cols, k, c_int_p))
    k += 10 * m
    L.orc_mass_diag(n, _ptr(np.ascontiguousarray(scene.mass), c_dbl_p), 0, at(vals, k, c_dbl_p),
                    at(rows, k, c_int_p), at(cols, k, c_int_p))
    return convert(vals, rows, cols)


def spmv(a, vals, rows, cols, x, b, y, n_rows):
    y = np.ascontiguousarray(y, dtype=np.float64).copy()
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_spmv(a, rows.size, _ptr(np.ascontiguousarray(vals), c_dbl_p), _ptr(rows, c_int_p),
                   _ptr(cols, c_int_p), _ptr(x, c_dbl_p), b, _ptr(y, c_dbl_p), n_rows)
    return y


# ---------------------------------------------------------------------- MAS
class Mas:
    """Oracle MAS preconditioner (identity slot map)."""

    def __init__(self, scene_or_n, bank_size=32, nbr=None, collision_node_offset=0):
        if nbr is None:
            s = scene_or_n
            n, nbr = s.n, (s.nbr_start, s.nbr_num, s.nbr_list)
            collision_node_offset = collision_node_offset or getattr(s, "n_abd_points", 0)
        else:
            n = int(scene_or_n)
        self.n, self.B = n, bank_size
        self._nbr = [np.ascontiguousarray(a, dtype=np.uint32) for a in nbr]
        if self._nbr[2].size == 0:
            self._nbr[2] = np.zeros(1, dtype=np.uint32)
        self.h = lib().orc_mas_create(n, bank_size, *[_ptr(a, c_uint_p) for a in self._nbr],
                                      collision_node_offset)
        if not self.h:
            raise ValueError("orc_mas_create failed")

    def build(self, vals, rows, cols, offset=0, pairs=None):
        pairs = np.zeros((0, 4), np.int32) if pairs is None else np.ascontiguousarray(pairs, np.int32)
        self._keep = (np.ascontiguousarray(vals), rows, cols, pairs)
        lib().orc_mas_build(self.h, rows.size, _ptr(self._keep[0], c_dbl_p), _ptr(rows, c_int_p),
                            _ptr(cols, c_int_p), offset, _ptr(pairs if pairs.size else None, c_int_p),
                            pairs.shape[0])

    def apply(self, r, fp32=False):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        lib().orc_mas_apply(self.h, _ptr(r, c_dbl_p), _ptr(z, c_dbl_p), int(bool(fp32)))
        return z

    def levels(self):
        L = ctypes.c_int()
        off = np.zeros(8, np.int32)
        cnt = np.zeros(8, np.int32)
        total = lib().orc_mas_levels(self.h, ctypes.byref(L), _ptr(off, c_int_p), _ptr(cnt, c_int_p))
        return L.value, off[:L.value + 1].copy(), cnt[:L.value].copy(), total

    def tables(self):
        L, off, cnt, total = self.levels()
        ct = np.zeros(max(self.n * (L - 1), 1), np.int32)
        gn = np.zeros(max(total, 1), np.int32)
        lib().orc_mas_tables(self.h, _ptr(ct, c_int_p), _ptr(gn, c_int_p))
        return ct[:self.n * (L - 1)].reshape(self.n, L - 1), gn[:total]

    def matrices(self, inverse=False):
        L, off, cnt, total = self.levels()
        nd = total // self.B
        nb = self.B * (self.B + 1) // 2
        out = np.zeros(nd * nb * 9)
        lib().orc_mas_matrices(self.h, _ptr(out, c_dbl_p), int(bool(inverse)))
        return out.reshape(nd, nb, 9)

    def close(self):
        if self.h:
            lib().orc_mas_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def num_levels(n, bank_size):
    return lib().orc_mas_num_levels_for(n, bank_size)


# ---------------------------------------------------------------------- PCG
@dataclasses.dataclass
class OracleSolve:
    x: np.ndarray
    iterations: int
    true_rel_residual: float
    seconds: float
    build_seconds: float


def pcg(vals, rows, cols, n_rows, b, mas=None, tol=1e-5, stop_mode=1, max_iter=None, mas_fp32=False):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    if max_iter is None:
        max_iter = max(2, int(0.3 * b.size))
    stats = np.zeros(4)
    lib().orc_pcg(rows.size, _ptr(np.ascontiguousarray(vals), c_dbl_p), _ptr(rows, c_int_p),
                  _ptr(cols, c_int_p), n_rows, mas.h if mas is not None else None, _ptr(b, c_dbl_p),
                  _ptr(x, c_dbl_p), tol, stop_mode, max_iter, int(bool(mas_fp32)), _ptr(stats, c_dbl_p))
    return x, int(stats[0]), float(stats[1])


class _Precond(ctypes.Structure):
    _fields_ = [("mas", ctypes.c_void_p), ("mas_row_offset", ctypes.c_int), ("mas_fp32", ctypes.c_int),
                ("n_bodies", ctypes.c_int), ("abd_inv", c_dbl_p)]


def pcg2(vals, rows, cols, n_rows, b, mas=None, mas_row_offset=0, abd_inv=None, tol=1e-5, stop_mode=1,
         max_iter=None, mas_fp32=False):
    """PCG with the subsystem preconditioner (block-Jacobi on the body rows,
    MAS on the rows from mas_row_offset)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    if max_iter is None:
        max_iter = max(2, int(0.3 * b.size))
    stats = np.zeros(4)
    inv = None if abd_inv is None else np.ascontiguousarray(abd_inv, dtype=np.float64).reshape(-1)
    pc = _Precond(mas.h if mas is not None else None, mas_row_offset, int(bool(mas_fp32)),
                  0 if abd_inv is None else int(abd_inv.shape[0]), _ptr(inv, c_dbl_p))
    lib().orc_pcg2(rows.size, _ptr(np.ascontiguousarray(vals), c_dbl_p), _ptr(rows, c_int_p), _ptr(cols, c_int_p),
                   n_rows, ctypes.byref(pc), _ptr(b, c_dbl_p), _ptr(x, c_dbl_p), tol, stop_mode, max_iter,
                   _ptr(stats, c_dbl_p))
    return x, int(stats[0]), float(stats[1])


def solve_scene(scene, bank_size=32, tol=1e-5, stop_mode=1, mas_fp32=False, max_iter=None):
    """Assemble, build the preconditioner and run PCG on the CPU (the plain
    FP64 path).  Scenes with affine bodies use block-Jacobi on the body rows
    and MAS on the FEM rows."""
    t0 = time.perf_counter()
    vals, rows, cols = assemble(scene)
    nb = getattr(scene, "n_bodies", 0)
    if nb:
        mas = Mas(scene, bank_size, collision_node_offset=scene.n_abd_points)
        mas.build(vals, rows, cols, 4 * nb, scene_pairs(scene))
        inv = abd_jacobi(vals, rows, cols, nb)
        t1 = time.perf_counter()
        x, it, res = pcg2(vals, rows, cols, scene.n_sys, scene.rhs, mas, 4 * nb, inv, tol, stop_mode, max_iter,
                          mas_fp32)
    else:
        mas = Mas(scene, bank_size)
        mas.build(vals, rows, cols, 0, scene_pairs(scene))
        t1 = time.perf_counter()
        x, it, res = pcg(vals, rows, cols, scene.n, scene.rhs, mas, tol, stop_mode, max_iter, mas_fp32)
    t2 = time.perf_counter()
    mas.close()
    return OracleSolve(x=x, iterations=it, true_rel_residual=res, seconds=t2 - t0, build_seconds=t1 - t0)


from . import cemas  # noqa: E402,F401  (CEMAS slot-map MAS, operator form)
