"""GPU parity of the affine-body (ABD) path through the C ABI against the
CPU oracle: body terms, the contact class mapping, the body block-Jacobi and
the whole hybrid solve.

Tolerances: body blocks <= 1e-10 relative per body (the device projects the
shape Hessian with its own Jacobi sweep order); contact blocks <= 1e-13
relative, rows/cols exact; block-Jacobi inverses <= 1e-9 relative; system
BCOO <= 1e-10; PCG iterate after a fixed iteration count <= 1e-8; the
converged solution's residual under the oracle's operator <= 1e-5.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p


def rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


@pytest.fixture(scope="module")
def hyb(pkg):
    return pkg.scenes.scene_hybrid_tiny(N=8, n_bodies=6, n_pairs=40)


@pytest.fixture(scope="module")
def hyb_big(pkg):
    return pkg.scenes.hybrid_scene("hybrid16", 16, 24, 600)


@pytest.mark.parametrize("which", ["tiny", "big"])
def test_abd_body_matches_oracle(pkg, orc, gpu_ctx, hyb, hyb_big, which):
    s = {"tiny": hyb, "big": hyb_big}[which]
    L, ctx = pkg._lib.lib(), gpu_ctx
    desc, keep = pkg.upload_scene(ctx, s)
    nb = s.n_bodies
    tv, tr, tc = ctx.empty((10 * nb, 9), np.float64), ctx.empty(10 * nb, np.int32), ctx.empty(10 * nb, np.int32)
    grad = ctx.empty(12 * nb, np.float64)
    pkg._lib.check(L.sgipc_abd_body_gradient_hessian(ctx.h, nb, vp(desc.q), vp(desc.q_tilde), vp(desc.mass10),
                                                      vp(desc.body_volume), s.kappa_abd, s.dt, vp(desc.body_fixed),
                                                      s.motor_speed, s.motor_strength, grad.vp, tv.vp, tr.vp, tc.vp), ctx.h)
    rv, rr, rc, rg = orc.abd_body(s)
    np.testing.assert_array_equal(tr.download(), rr)
    np.testing.assert_array_equal(tc.download(), rc)
    gv = tv.download()
    err = np.abs(gv - rv).reshape(nb, -1).max(1)
    mag = np.abs(rv).reshape(nb, -1).max(1)
    assert np.all(err <= 1e-10 * mag), float((err / mag).max())
    assert rel(grad.download(), rg) <= 1e-10


def test_contact_partition_matches_oracle(pkg, orc, gpu_ctx, hyb_big):
    s = hyb_big
    L, ctx = pkg._lib.lib(), gpu_ctx
    bv, br, bc, _ = orc.barrier_pt(s, with_grad=False)
    # mix in body-body and same-body blocks by relabelling some FEM points as body points
    rng = np.random.default_rng(3)
    sel = rng.random(br.size) < 0.2
    br = br.copy()
    bc = bc.copy()
    br[sel] = rng.integers(0, s.n_abd_points, sel.sum())
    bc[sel] = np.maximum(br[sel], rng.integers(0, s.n_abd_points, sel.sum()))
    ov, orr, oc = orc.abd_contact_map(s, bv, br, bc)
    dv, dr, dc = ctx.to_device(bv, np.float64), ctx.to_device(br, np.int32), ctx.to_device(bc, np.int32)
    cap = 16 * br.size
    gv, gr, gc = ctx.empty((cap, 9), np.float64), ctx.empty(cap, np.int32), ctx.empty(cap, np.int32)
    desc, keep = pkg.upload_scene(ctx, s)
    n_out = ctypes.c_int()
    pkg._lib.check(L.sgipc_contact_partition(ctx.h, dv.vp, dr.vp, dc.vp, br.size, s.n_abd_points,
                                              vp(desc.point_body), vp(desc.x_bar), s.n_bodies, vp(desc.body_fixed),
                                              vp(desc.boundary_type), gv.vp, gr.vp, gc.vp, cap, ctypes.byref(n_out)), ctx.h)
    m = n_out.value
    assert m == orr.size
    np.testing.assert_array_equal(gr.download()[:m], orr)
    np.testing.assert_array_equal(gc.download()[:m], oc)
    assert rel(gv.download()[:m], ov) <= 1e-13


def test_abd_jacobi_matches_oracle(pkg, orc, gpu_ctx, hyb_big):
    s = hyb_big
    L, ctx = pkg._lib.lib(), gpu_ctx
    vals, rows, cols = orc.assemble(s)
    nb = s.n_bodies
    dv, dr, dc = ctx.to_device(vals, np.float64), ctx.to_device(rows, np.int32), ctx.to_device(cols, np.int32)
    inv = ctx.empty((nb, 144), np.float64)
    pkg._lib.check(L.sgipc_abd_jacobi_build(ctx.h, dv.vp, dr.vp, dc.vp, rows.size, nb, inv.vp), ctx.h)
    ref = orc.abd_jacobi(vals, rows, cols, nb)
    got = inv.download().reshape(nb, 12, 12)
    for b in range(nb):
        assert rel(got[b], ref[b]) <= 1e-9, b
    r = np.random.default_rng(2).standard_normal(12 * nb)
    dr_ = ctx.to_device(r)
    dz = ctx.empty(12 * nb, np.float64)
    pkg._lib.check(L.sgipc_abd_jacobi_apply(ctx.h, inv.vp, nb, dr_.vp, dz.vp), ctx.h)
    zr = np.einsum("bij,bj->bi", ref, r.reshape(nb, 12)).reshape(-1)
    assert rel(dz.download(), zr) <= 1e-9


@pytest.mark.parametrize("which", ["tiny", "big"])
def test_hybrid_system_solve_matches_oracle(pkg, orc, hyb, hyb_big, which):
    s = {"tiny": hyb, "big": hyb_big}[which]
    ls = pkg.LinearSystem(s, bank_size=32)
    try:
        res = ls.solve(tol=1e-5)
        assert res.status == 0 and res.true_rel_residual <= 1e-5, res
        gv, gr, gc = ls.sys.bcoo()
        ov, orr, oc = orc.assemble(s)
        np.testing.assert_array_equal(gr, orr)
        np.testing.assert_array_equal(gc, oc)
        assert rel(gv, ov) <= 1e-10
        ax = orc.spmv(1.0, ov, orr, oc, res.x, 0.0, np.zeros(3 * s.n_sys), s.n_sys)
        assert np.linalg.norm(s.rhs - ax) / np.linalg.norm(s.rhs) <= 1.01e-5
        for k in (3, 12):
            fixed = ls.solve(tol=1e-30, max_iter=k)
            xr = orc.solve_scene(s, 32, tol=1e-30, max_iter=k).x
            assert np.linalg.norm(fixed.x - xr) / np.linalg.norm(xr) <= 1e-8, k
    finally:
        ls.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_hybrid_two_grid_cycle(pkg, orc, hyb_big, precision):
    """Coarse mode 2 on a hybrid system: the MAS covers the rows past the
    body rows, so it builds A P and its own fine operator from the BCOO's
    suffix of those rows (renumbered from 0).  The solve takes the two-grid
    path, meets the tolerance under the oracle's product, and (FP64) agrees
    with the level-1 cycle's solution of the same system."""
    s = hyb_big
    ls = pkg.LinearSystem(s, bank_size=32, coarse_mode=2)
    m1 = None
    try:
        res = ls.solve(tol=1e-5, precision=precision)
        assert res.status == 0 and res.coarse_mode == 2 and res.level2_exact, res
        ov, orr, oc = orc.assemble(s)
        ax = orc.spmv(1.0, ov, orr, oc, res.x, 0.0, np.zeros(3 * s.n_sys), s.n_sys)
        assert np.linalg.norm(s.rhs - ax) / np.linalg.norm(s.rhs) <= 1.01e-5
        if precision == "fp64":
            tight = ls.solve(tol=1e-9)
            assert tight.status == 0 and tight.coarse_mode == 2
            m1 = pkg.LinearSystem(s, bank_size=32, coarse_mode=1)
            ref = m1.solve(tol=1e-9)
            assert ref.status == 0 and ref.coarse_mode == 1
            assert np.linalg.norm(tight.x - ref.x) / np.linalg.norm(ref.x) <= 1e-4
    finally:
        if m1 is not None:
            m1.close()
        ls.close()


# ---- Dirichlet vertices, the assembled right-hand side, motor bodies
@pytest.fixture(scope="module")
def hyb_bc(pkg):
    """Hybrid scene with the grid's bottom layer pinned (Dirichlet), two
    fixed bodies and two motor bodies."""
    s = pkg.scenes.hybrid_scene("hybrid12_bc", 12, 12, 300, n_fixed=2, n_motor=2)
    return pkg.scenes.with_dirichlet(s)


def test_motor_bodies_match_oracle(pkg, orc, gpu_ctx, hyb_bc):
    """Body terms with motor bodies (BodyBoundaryType 2): the device's closed
    form of the motor target against the oracle's J^-1 J_delta restatement."""
    s = hyb_bc
    L, ctx = pkg._lib.lib(), gpu_ctx
    desc, keep = pkg.upload_scene(ctx, s)
    nb = s.n_bodies
    assert (s.body_fixed == 2).sum() == 2
    tv, tr, tc = ctx.empty((10 * nb, 9), np.float64), ctx.empty(10 * nb, np.int32), ctx.empty(10 * nb, np.int32)
    grad = ctx.empty(12 * nb, np.float64)
    pkg._lib.check(L.sgipc_abd_body_gradient_hessian(ctx.h, nb, vp(desc.q), vp(desc.q_tilde), vp(desc.mass10),
                                                      vp(desc.body_volume), s.kappa_abd, s.dt, vp(desc.body_fixed),
                                                      s.motor_speed, s.motor_strength, grad.vp, tv.vp, tr.vp, tc.vp),
                   ctx.h)
    rv, rr, rc, rg = orc.abd_body(s)
    gv = tv.download()
    err = np.abs(gv - rv).reshape(nb, -1).max(1)
    mag = np.abs(rv).reshape(nb, -1).max(1)
    assert np.all(err <= 1e-10 * mag), float((err / mag).max())
    g = grad.download().reshape(nb, 12)
    rg = rg.reshape(nb, 12)
    for b in range(nb):
        assert np.abs(g[b] - rg[b]).max() <= 1e-10 * max(np.abs(rg[b]).max(), 1e-300), b
    # the motor term is present: the motor bodies differ from the same bodies set free
    free = pkg.scenes.hybrid_scene("hybrid12_bc", 12, 12, 300, n_fixed=2, n_motor=0)
    _, _, _, fg = orc.abd_body(free)
    m = s.body_fixed == 2
    assert np.abs(fg.reshape(nb, 12)[m] - rg[m]).max() > 0


def test_dirichlet_system_matches_oracle(pkg, orc, hyb_bc):
    """Whole hybrid solve with Dirichlet vertices: the device BCOO (blocks
    touching a pinned vertex zeroed, mass kept) and the assembled gradient
    (bodies: M(q - q~) + shape + J^T contact gradient; FEM rows zero at pinned
    vertices) equal the oracle's; solving H dx = gradient leaves the pinned
    vertices decoupled: their rows hold only the mass block, so
    m |dx| <= ||b - A dx|| there (the MAS coarse levels still spread into
    them, as in the reference, whose step then leaves pinned vertices alone)."""
    s = hyb_bc
    ls = pkg.LinearSystem(s, bank_size=32)
    try:
        ov, orr, oc = orc.assemble(s)
        rv, rr, rc, rg = orc.triplets(s, with_grad=True)
        res = ls.solve(tol=1e-5, want_grad=True)
        gv, gr, gc = ls.sys.bcoo()
        np.testing.assert_array_equal(gr, orr)
        np.testing.assert_array_equal(gc, oc)
        assert rel(gv, ov) <= 1e-10
        g = ls.gradient()
        assert rel(g, rg) <= 1e-10, rel(g, rg)
        pinned = np.nonzero(s.boundary_type)[0]
        off = 4 * s.n_bodies
        assert np.all(g[3 * off:].reshape(-1, 3)[pinned] == 0.0)
        # pinned rows and columns carry only the mass block
        bad = ((gr >= off) & np.isin(gr - off, pinned)) | ((gc >= off) & np.isin(gc - off, pinned))
        offdiag = bad & (gr != gc)
        assert np.all(gv[offdiag] == 0.0)
        # the Newton system with the assembled gradient as right-hand side
        ls.sys.b.upload(rg)
        res = ls.solve(tol=1e-5)
        assert res.status == 0 and res.true_rel_residual <= 1e-5
        ax = orc.spmv(1.0, ov, orr, oc, res.x, 0.0, np.zeros(3 * s.n_sys), s.n_sys)
        rr_ = rg - ax
        assert np.linalg.norm(rr_) / np.linalg.norm(rg) <= 1.01e-5
        xd = res.x[3 * off:].reshape(-1, 3)[pinned]
        np.testing.assert_allclose(s.mass[pinned, None] * xd, -rr_[3 * off:].reshape(-1, 3)[pinned],
                                   rtol=1e-9, atol=1e-300)
    finally:
        ls.close()


def test_fem_dirichlet_gradient_matches_oracle(pkg, orc):
    """FEM-only scene with pinned bottom vertices and contacts: BCOO and
    gradient against the oracle, FP32 refinement with the two-grid cycle."""
    s = pkg.scenes.with_dirichlet(pkg.scenes.tet_grid_scene("bc_grid12", 12, E_loguniform=(1e4, 1e9), n_pairs=300))
    ls = pkg.LinearSystem(s, bank_size=32, coarse_mode=2)
    try:
        rv, rr, rc, rg = orc.triplets(s, with_grad=True)
        ov, orr, oc = orc.assemble(s)
        res = ls.solve(tol=1e-5, want_grad=True, precision="fp32_ir")
        gv, gr, gc = ls.sys.bcoo()
        np.testing.assert_array_equal(gr, orr)
        assert rel(gv, ov) <= 1e-10
        assert rel(ls.gradient(), rg) <= 1e-10
        ls.sys.b.upload(rg)
        res = ls.solve(tol=1e-5, precision="fp32_ir")
        assert res.status == 0 and res.true_rel_residual <= 1e-5
        pinned = np.nonzero(s.boundary_type)[0]
        ax = orc.spmv(1.0, ov, orr, oc, res.x, 0.0, np.zeros(3 * s.n), s.n)
        # decoupled pinned rows: m dx = -(b - A dx) there
        np.testing.assert_allclose(s.mass[pinned, None] * res.x.reshape(-1, 3)[pinned],
                                   -(rg - ax).reshape(-1, 3)[pinned], rtol=1e-9, atol=1e-300)
    finally:
        ls.close()
