"""Failure paths through the C ABI: each status code the ABI documents is
returned to the caller, with the message in sgipc_last_error, and nothing
aborts.  Each condition is forced from the host (sizes, capacities, an
indefinite matrix, an iteration cap), so every kernel exits normally."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p


@pytest.fixture(scope="module")
def _lib(pkg):
    return pkg._lib


def upload(ctx, vals, rows, cols):
    return ctx.to_device(vals, np.float64), ctx.to_device(rows, np.int32), ctx.to_device(cols, np.int32)


def test_empty_inputs_return_empty(pkg, _lib, gpu_ctx):
    L, h = _lib.lib(), gpu_ctx.h
    d = gpu_ctx.empty(9, np.float64)
    i = gpu_ctx.empty(1, np.int32)
    out = ctypes.c_int(-1)
    assert L.sgipc_convert_srbk(h, d.vp, i.vp, i.vp, 0, d.vp, i.vp, i.vp, 1, ctypes.byref(out)) == _lib.SGIPC_EMPTY
    assert out.value == 0
    assert L.sgipc_spmv_srbk(h, 1.0, d.vp, i.vp, i.vp, 1, d.vp, 0.0, d.vp, 0) == _lib.SGIPC_EMPTY
    opts = _lib.PcgOpts()
    L.sgipc_pcg_default_opts(ctypes.byref(opts))
    st = _lib.PcgStats()
    assert L.sgipc_pcg_solve(h, d.vp, i.vp, i.vp, 0, 1, None, d.vp, d.vp, ctypes.byref(opts),
                             ctypes.byref(st)) == _lib.SGIPC_EMPTY
    assert L.sgipc_pcg_solve(h, d.vp, i.vp, i.vp, 1, 0, None, d.vp, d.vp, ctypes.byref(opts),
                             ctypes.byref(st)) == _lib.SGIPC_EMPTY
    assert L.sgipc_fem_tet_gradient_hessian(h, d.vp, i.vp, d.vp, d.vp, d.vp, d.vp, 0, 0.01, 0, None, d.vp, i.vp,
                                            i.vp) == _lib.SGIPC_EMPTY
    sysh = vp()
    desc = _lib.SceneDesc()
    assert L.sgipc_system_create(h, ctypes.byref(sysh), ctypes.byref(desc), 32) == _lib.SGIPC_EMPTY
    assert not sysh.value


def test_iteration_cap_returns_maxiter(pkg, _lib, orc, gpu_ctx, scene_c0):
    s = scene_c0
    vals, rows, cols = orc.assemble(s)
    dv, dr, dc = upload(gpu_ctx, vals, rows, cols)
    db = gpu_ctx.to_device(s.rhs)
    dx = gpu_ctx.empty(3 * s.n, np.float64)
    opts = _lib.PcgOpts()
    _lib.lib().sgipc_pcg_default_opts(ctypes.byref(opts))
    opts.max_iter = 2
    st = _lib.PcgStats()
    status = _lib.lib().sgipc_pcg_solve(gpu_ctx.h, dv.vp, dr.vp, dc.vp, rows.size, s.n, None, db.vp, dx.vp,
                                        ctypes.byref(opts), ctypes.byref(st))
    assert status == _lib.SGIPC_MAXITER == st.status
    # the iterate it stopped at is written and finite, and its residual reported
    assert st.iterations == 1 and 0 < st.rel_residual < 1
    assert np.all(np.isfinite(dx.download()))


def test_undersized_outputs_return_capacity(pkg, _lib, orc, gpu_ctx, scene_tiny):
    L, h = _lib.lib(), gpu_ctx.h
    rv, rr, rc, _ = orc.triplets(scene_tiny)
    ov, orr, oc = orc.convert(rv, rr, rc)
    tv, tr, tc = upload(gpu_ctx, rv, rr, rc)
    cap = orr.size - 1
    ovd, ord_, ocd = gpu_ctx.empty((cap, 9), np.float64), gpu_ctx.empty(cap, np.int32), gpu_ctx.empty(cap, np.int32)
    out = ctypes.c_int(0)
    st = L.sgipc_convert_srbk(h, tv.vp, tr.vp, tc.vp, rr.size, ovd.vp, ord_.vp, ocd.vp, cap, ctypes.byref(out))
    assert st == _lib.SGIPC_CAPACITY, st
    assert b"capacity" in L.sgipc_last_error(h)
    assert out.value == orr.size  # the count the caller must provide


def test_contact_partition_capacity(pkg, _lib, orc, gpu_ctx):
    L, h = _lib.lib(), gpu_ctx.h
    s = pkg.scenes.scene_hybrid_tiny(N=6, n_bodies=4, n_pairs=10)
    desc, keep = pkg.upload_scene(gpu_ctx, s)
    p = s.n_pairs
    cv, cr, cc = gpu_ctx.empty((16 * p, 9), np.float64), gpu_ctx.empty(16 * p, np.int32), \
        gpu_ctx.empty(16 * p, np.int32)
    _lib.check(L.sgipc_barrier_pt_gradient_hessian(h, None, vp(desc.pairs), vp(desc.pair_x), p, s.dhat ** 2,
                                                    s.kappa, 0, None, cv.vp, cr.vp, cc.vp), h)
    need = ctypes.c_int(0)
    big = 64 * p
    ov, orr, oc = gpu_ctx.empty((big, 9), np.float64), gpu_ctx.empty(big, np.int32), gpu_ctx.empty(big, np.int32)
    _lib.check(L.sgipc_contact_partition(h, cv.vp, cr.vp, cc.vp, 16 * p, s.n_abd_points, vp(desc.point_body),
                                         vp(desc.x_bar), s.n_bodies, vp(desc.body_fixed), None, ov.vp, orr.vp, oc.vp, big,
                                         ctypes.byref(need)), h)
    assert 0 < need.value <= big
    got = ctypes.c_int(0)
    st = L.sgipc_contact_partition(h, cv.vp, cr.vp, cc.vp, 16 * p, s.n_abd_points, vp(desc.point_body),
                                   vp(desc.x_bar), s.n_bodies, vp(desc.body_fixed), None, ov.vp, orr.vp, oc.vp,
                                   need.value - 1, ctypes.byref(got))
    assert st == _lib.SGIPC_CAPACITY, st
    assert got.value == need.value  # the count the caller must provide


def test_indefinite_matrix_returns_not_spd(pkg, _lib, gpu_ctx):
    """Diagonal blocks diag(1, -1, 1) and b = (1, 2, 0) per vertex: p.Ap =
    1 - 4 < 0 at the first iteration (no preconditioner)."""
    n = 64
    vals = np.zeros((n, 9))
    vals[:, 0], vals[:, 4], vals[:, 8] = 1.0, -1.0, 1.0
    rows = np.arange(n, dtype=np.int32)
    dv, dr, dc = upload(gpu_ctx, vals, rows, rows.copy())
    b = np.zeros(3 * n)
    b[0::3], b[1::3] = 1.0, 2.0
    db = gpu_ctx.to_device(b)
    dx = gpu_ctx.empty(3 * n, np.float64)
    with pytest.raises(pkg.SgipcError) as ei:
        pkg.PCGSolver(gpu_ctx, tol=1e-5).solve(dv, dr, dc, n, n, None, db, dx)
    assert ei.value.status == _lib.SGIPC_NOT_SPD
