"""GPU parity tests: every HIP kernel, called through the C ABI, against the
CPU oracle on the same seeded inputs.

Tolerances (floating point; integer/index work is compared bit-exactly):
  FP64 SpMV               max|y_gpu - y_orc| <= 1e-13 * max|y_orc|   (summation order)
  FP32 SpMV               <= 2e-6 relative
  triplet -> BCOO         bit-exact (both sum each block in arena order)
  tet / barrier Hessians  max|dH| <= 1e-10 * max|H| per element; rows/cols exact
  MAS hierarchy           exact (levels, counts, offsets, coarse tables)
  MAS inverse (FP64)      <= 1e-9 relative per domain
  MAS apply               FP64 <= 1e-9 (the inverse's own bound: both sides
                          run Gauss-Jordan on blocks with condition ~1e8 in a
                          different operation order), FP32 <= 2e-5 relative
  PCG iterate x_k         <= 1e-8 relative after a fixed number of iterations
  full solve              true FP64 relative residual <= 1e-5; the iterate after a
                          fixed iteration count within 1e-8 of the oracle's; on
                          the well-conditioned 16^3 scene, tight (1e-11) solves
                          within 1e-8 relative
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def upload_bcoo(ctx, vals, rows, cols):
    return (ctx.to_device(vals, np.float64), ctx.to_device(rows, np.int32), ctx.to_device(cols, np.int32))


# --------------------------------------------------------------------- SpMV
@pytest.mark.parametrize("which", ["tiny", "small", "c0"])
def test_spmv_f64_matches_oracle(pkg, orc, gpu_ctx, which, scene_tiny, scene_small, scene_c0):
    s = {"tiny": scene_tiny, "small": scene_small, "c0": scene_c0}[which]
    vals, rows, cols = orc.assemble(s)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(3 * s.n)
    y0 = rng.standard_normal(3 * s.n)
    ref = orc.spmv(1.3, vals, rows, cols, x, 0.7, y0, s.n)
    dv, dr, dc = upload_bcoo(gpu_ctx, vals, rows, cols)
    dx = gpu_ctx.to_device(x)
    dy = gpu_ctx.to_device(y0)
    pkg.Spmv(gpu_ctx).srbk_spmv(1.3, dv, dr, dc, rows.size, dx, 0.7, dy, s.n)
    got = dy.download()
    assert rel(got, ref) <= 1e-13, rel(got, ref)
    # b = 0 overwrites y
    dy.upload(np.full(3 * s.n, np.nan))
    pkg.Spmv(gpu_ctx).srbk_spmv(1.0, dv, dr, dc, rows.size, dx, 0.0, dy, s.n)
    ref0 = orc.spmv(1.0, vals, rows, cols, x, 0.0, np.zeros(3 * s.n), s.n)
    assert rel(dy.download(), ref0) <= 1e-13


@pytest.mark.parametrize("form", ["halo", "gather"])
@pytest.mark.parametrize("which", ["tiny", "small", "c0", "cloth_tiny", "cloth64", "grid24_contact"])
def test_tiled_spmv_matches_oracle(pkg, orc, gpu_ctx, which, form, scene_tiny, scene_small, scene_c0, monkeypatch):
    """The PCG loop's row-tiled SpMV (one tile for the tiny scenes, several
    with a partly filled last tile for the others), FP64, FP32 and FP32
    values with FP64 vectors; both layout forms: column words as indices of
    the tile's LDS copy of x (halo), and as global columns (gather)."""
    s = {"tiny": scene_tiny, "small": scene_small, "c0": scene_c0}.get(which)
    if s is None:
        if which == "cloth_tiny":
            s = pkg.scenes.scene_cloth_tiny(N=12, n_pairs=30)
        elif which == "cloth64":
            s = pkg.scenes.cloth_grid_scene("cloth64", 64, n_pairs=400)
        else:  # 27 tiles, contact pairs across the whole grid: large halos
            s = pkg.scenes.tet_grid_scene("grid24_contact", 24, E_loguniform=(1e4, 1e9), n_pairs=3000)
    vals, rows, cols = orc.assemble(s)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(3 * s.n)
    y0 = rng.standard_normal(3 * s.n)
    dv, dr, dc = upload_bcoo(gpu_ctx, vals, rows, cols)
    if form == "gather":
        monkeypatch.setenv("SGIPC_SPMV_GATHER", "1")
    A = pkg.SpMat(gpu_ctx, dv, dr, dc, rows.size, s.n, want_f32=True)
    monkeypatch.delenv("SGIPC_SPMV_GATHER", raising=False)
    try:
        halo_form, halo_max, tl_slots = A.layout()
        assert halo_form == (1 if form == "halo" else 0)
        assert tl_slots >= 1 and halo_max >= 0
        if s.n <= 512:
            assert halo_max == 0  # one tile: nothing outside it
        dx, dy = gpu_ctx.to_device(x), gpu_ctx.to_device(y0)
        A.apply(0.9, dx, -0.4, dy)
        assert rel(dy.download(), orc.spmv(0.9, vals, rows, cols, x, -0.4, y0, s.n)) <= 1e-13
        dy.upload(np.full(3 * s.n, np.nan))
        A.apply(1.0, dx, 0.0, dy)
        ref = orc.spmv(1.0, vals, rows, cols, x, 0.0, np.zeros(3 * s.n), s.n)
        assert rel(dy.download(), ref) <= 1e-13
        x32 = gpu_ctx.to_device(x.astype(np.float32), np.float32)
        y32 = gpu_ctx.to_device(np.zeros(3 * s.n, np.float32), np.float32)
        A.apply(1.0, x32, 0.0, y32)
        assert rel(y32.download().astype(np.float64), ref) <= 2e-6
        # FP32 values, FP64 vectors: the matrix, the LDS copy of x and the in-tile
        # transposed products rounded to FP32 (~6e-8 relative each)
        dy.upload(y0)
        A.apply_f32_values(0.9, dx, -0.4, dy)
        assert rel(dy.download(), orc.spmv(0.9, vals, rows, cols, x, -0.4, y0, s.n)) <= 2e-6
    finally:
        A.close()


def test_spmv_f32_matches_oracle(pkg, orc, gpu_ctx, scene_c0):
    s = scene_c0
    vals, rows, cols = orc.assemble(s)
    x = np.random.default_rng(3).standard_normal(3 * s.n)
    v32 = vals.astype(np.float32)
    x32 = x.astype(np.float32)
    ref = orc.spmv(1.0, v32.astype(np.float64), rows, cols, x32.astype(np.float64), 0.0, np.zeros(3 * s.n), s.n)
    dv = gpu_ctx.to_device(v32)
    dr = gpu_ctx.to_device(rows)
    dc = gpu_ctx.to_device(cols)
    dx = gpu_ctx.to_device(x32)
    dy = gpu_ctx.empty(3 * s.n, np.float32)
    pkg.Spmv(gpu_ctx).srbk_spmv(1.0, dv, dr, dc, rows.size, dx, 0.0, dy, s.n)
    assert rel(dy.download().astype(np.float64), ref) <= 2e-6


def test_spmv_symmetry_identity(pkg, orc, gpu_ctx, scene_small):
    """u.(A v) == v.(A u) for the symmetric operator (dropping the transposed
    part or transposing a block breaks it)."""
    s = scene_small
    vals, rows, cols = orc.assemble(s)
    dv, dr, dc = upload_bcoo(gpu_ctx, vals, rows, cols)
    rng = np.random.default_rng(11)
    u, v = rng.standard_normal(3 * s.n), rng.standard_normal(3 * s.n)
    out = []
    for a, b in ((u, v), (v, u)):
        dx = gpu_ctx.to_device(b)
        dy = gpu_ctx.empty(3 * s.n, np.float64)
        pkg.Spmv(gpu_ctx).srbk_spmv(1.0, dv, dr, dc, rows.size, dx, 0.0, dy, s.n)
        out.append(float(a @ dy.download()))
    assert abs(out[0] - out[1]) <= 1e-12 * max(abs(out[0]), 1.0)


# ---------------------------------------------------------------- assembly
def _device_triplets(pkg, ctx, s):
    from stiff_gipc_amd import _lib
    L = _lib.lib()
    m, p, n = s.n_tets, s.n_pairs, s.n
    nt = 16 * p + 10 * m + n
    tv = ctx.empty((nt, 9), np.float64)
    tr = ctx.empty(nt, np.int32)
    tc = ctx.empty(nt, np.int32)
    grad = ctx.empty(3 * n, np.float64).zero()
    desc, keep = pkg.upload_scene(ctx, s)
    import ctypes
    vp = ctypes.c_void_p
    if p:
        _lib.check(L.sgipc_barrier_pt_gradient_hessian(ctx.h, vp(desc.x), vp(desc.pairs), vp(desc.pair_x), p,
                                                        s.dhat ** 2, s.kappa, 0, grad.vp, tv.vp, tr.vp, tc.vp), ctx.h)
    off = 16 * p
    _lib.check(L.sgipc_fem_tet_gradient_hessian(ctx.h, vp(desc.x), vp(desc.tets), vp(desc.dm_inv), vp(desc.vol),
                                                 vp(desc.mu), vp(desc.lam), m, s.dt, 0, grad.vp,
                                                 vp(tv.ptr + 72 * off), vp(tr.ptr + 4 * off), vp(tc.ptr + 4 * off)),
               ctx.h)
    off += 10 * m
    _lib.check(L.sgipc_mass_diagonal(ctx.h, vp(desc.mass), n, 0, vp(tv.ptr + 72 * off), vp(tr.ptr + 4 * off),
                                      vp(tc.ptr + 4 * off)), ctx.h)
    return tv, tr, tc, grad, keep


@pytest.mark.parametrize("which", ["tiny", "small", "c0"])
def test_element_hessians_match_oracle(pkg, orc, gpu_ctx, which, scene_tiny, scene_small, scene_c0):
    s = {"tiny": scene_tiny, "small": scene_small, "c0": scene_c0}[which]
    tv, tr, tc, grad, keep = _device_triplets(pkg, gpu_ctx, s)
    rv, rr, rc, rg = orc.triplets(s, with_grad=True)
    gv, gr, gc = tv.download(), tr.download(), tc.download()
    np.testing.assert_array_equal(gr, rr)
    np.testing.assert_array_equal(gc, rc)
    p16 = 16 * s.n_pairs
    # barrier blocks
    if s.n_pairs:
        b_err = np.abs(gv[:p16] - rv[:p16]).reshape(s.n_pairs, -1).max(1)
        b_mag = np.abs(rv[:p16]).reshape(s.n_pairs, -1).max(1)
        assert np.all(b_err <= 1e-10 * b_mag + 1e-300), float((b_err / b_mag).max())
    # tet blocks, per element
    t = slice(p16, p16 + 10 * s.n_tets)
    e_err = np.abs(gv[t] - rv[t]).reshape(s.n_tets, -1).max(1)
    e_mag = np.abs(rv[t]).reshape(s.n_tets, -1).max(1)
    assert np.all(e_err <= 1e-10 * e_mag), float((e_err / e_mag).max())
    np.testing.assert_array_equal(gv[p16 + 10 * s.n_tets:], rv[p16 + 10 * s.n_tets:])
    g = grad.download()
    assert rel(g, rg) <= 1e-10, rel(g, rg)


@pytest.mark.parametrize("which", ["tiny", "c0"])
def test_convert_bit_exact(pkg, orc, gpu_ctx, which, scene_tiny, scene_c0):
    s = {"tiny": scene_tiny, "c0": scene_c0}[which]
    rv, rr, rc, _ = orc.triplets(s)
    ov, orr, oc = orc.convert(rv, rr, rc)
    dv, dr, dc = upload_bcoo(gpu_ctx, rv, rr, rc)
    gv, gr, gc, nnz = pkg.Converter(gpu_ctx).srbk_convert(dv, dr, dc, rr.size)
    assert nnz == orr.size
    np.testing.assert_array_equal(gr.download()[:nnz], orr)
    np.testing.assert_array_equal(gc.download()[:nnz], oc)
    np.testing.assert_array_equal(gv.download()[:nnz], ov)


def test_convert_rejects_lower_triangle(pkg, gpu_ctx):
    from stiff_gipc_amd import SgipcError
    from stiff_gipc_amd._lib import SGIPC_BAD_INDEX
    v = gpu_ctx.to_device(np.ones((2, 9)))
    r = gpu_ctx.to_device(np.array([0, 3], np.int32))
    c = gpu_ctx.to_device(np.array([1, 2], np.int32))
    with pytest.raises(SgipcError) as ei:
        pkg.Converter(gpu_ctx).srbk_convert(v, r, c, 2)
    assert ei.value.status == SGIPC_BAD_INDEX


# --------------------------------------------------------------------- MAS
def _gpu_mas(pkg, ctx, s, B, vals, rows, cols, precision):
    nbr = [ctx.to_device(a, np.uint32) for a in (s.nbr_start, s.nbr_num,
                                                 s.nbr_list if s.nbr_list.size else np.zeros(1, np.uint32))]
    mas = pkg.MASPreconditioner(ctx, B)
    mas.initPreconditioner_Neighbor(s.n, 0, int(s.nbr_list.size), *nbr)
    mas.initPreconditioner_Matrix()
    dv, dr, dc = upload_bcoo(ctx, vals, rows, cols)
    pairs = ctx.to_device(s.pairs.reshape(-1), np.int32) if s.n_pairs else None
    mas.setPreconditioner_bcoo(dv, dr, dc, rows.size, 0, pairs, s.n_pairs, precision)
    return mas, (nbr, dv, dr, dc, pairs)


@pytest.mark.parametrize("B", [16, 32])
@pytest.mark.parametrize("which", ["tiny", "small", "c0"])
def test_mas_hierarchy_and_inverse_match_oracle(pkg, orc, gpu_ctx, B, which, scene_tiny, scene_small, scene_c0):
    s = {"tiny": scene_tiny, "small": scene_small, "c0": scene_c0}[which]
    vals, rows, cols = orc.assemble(s)
    om = orc.Mas(s, B)
    om.build(vals, rows, cols, 0, s.pairs)
    mas, keep = _gpu_mas(pkg, gpu_ctx, s, B, vals, rows, cols, pkg.FP64)
    L, off, cnt, total = mas.levels()
    oL, ooff, ocnt, ototal = om.levels()
    assert L == oL and total == ototal
    np.testing.assert_array_equal(off, ooff)
    np.testing.assert_array_equal(cnt, ocnt)
    ct, gn = mas.tables()
    oct_, ogn = om.tables()
    np.testing.assert_array_equal(ct, oct_)
    # going_next is defined for the real nodes of every level
    for l in range(L):
        seg = slice(off[l], off[l] + cnt[l])
        np.testing.assert_array_equal(gn[seg], ogn[seg])
    inv = mas.inverse_blocks()
    oinv = om.matrices(inverse=True)
    for d in range(inv.shape[0]):
        assert rel(inv[d], oinv[d]) <= 1e-9, (d, rel(inv[d], oinv[d]))
    mas.close()
    om.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("which", ["tiny", "small", "c0"])
def test_mas_apply_matches_oracle(pkg, orc, gpu_ctx, precision, which, scene_tiny, scene_small, scene_c0):
    s = {"tiny": scene_tiny, "small": scene_small, "c0": scene_c0}[which]
    B = 32
    vals, rows, cols = orc.assemble(s)
    om = orc.Mas(s, B)
    om.build(vals, rows, cols, 0, s.pairs)
    prec = pkg.FP64 if precision == "fp64" else pkg.FP32_MAS
    mas, keep = _gpu_mas(pkg, gpu_ctx, s, B, vals, rows, cols, prec)
    r = np.random.default_rng(5).standard_normal(3 * s.n)
    dr_ = gpu_ctx.to_device(r)
    dz = gpu_ctx.empty(3 * s.n, np.float64)
    mas.preconditioning(dr_, dz)
    z = dz.download()
    zr = om.apply(r, fp32=(precision == "fp32"))
    tol = 1e-9 if precision == "fp64" else 2e-5
    assert rel(z, zr) <= tol, rel(z, zr)
    mas.close()
    om.close()


# --------------------------------------------------------------------- PCG
@pytest.mark.parametrize("which", ["small", "c0"])
def test_pcg_fixed_iterations_match_oracle(pkg, orc, gpu_ctx, which, scene_small, scene_c0):
    s = {"small": scene_small, "c0": scene_c0}[which]
    vals, rows, cols = orc.assemble(s)
    om = orc.Mas(s, 32)
    om.build(vals, rows, cols, 0, s.pairs)
    k = 12
    xr, itr, _ = orc.pcg(vals, rows, cols, s.n, s.rhs, om, tol=1e-30, stop_mode=1, max_iter=k)
    mas, keep = _gpu_mas(pkg, gpu_ctx, s, 32, vals, rows, cols, pkg.FP64)
    dv, dr, dc = keep[1], keep[2], keep[3]
    db = gpu_ctx.to_device(s.rhs)
    dx = gpu_ctx.empty(3 * s.n, np.float64)
    st = pkg.PCGSolver(gpu_ctx, tol=1e-30, max_iter=k).solve(dv, dr, dc, rows.size, s.n, mas, db, dx)
    assert st.iterations == itr == k - 1
    assert rel(dx.download(), xr) <= 1e-8
    mas.close()
    om.close()


@pytest.mark.parametrize("stop_mode", [0, 1])
def test_pcg_stop_rules_match_oracle(pkg, orc, gpu_ctx, scene_c0, stop_mode):
    s = scene_c0
    vals, rows, cols = orc.assemble(s)
    om = orc.Mas(s, 32)
    om.build(vals, rows, cols, 0, s.pairs)
    tol = 1e-4 if stop_mode == 0 else 1e-5
    xr, itr, resr = orc.pcg(vals, rows, cols, s.n, s.rhs, om, tol=tol, stop_mode=stop_mode)
    mas, keep = _gpu_mas(pkg, gpu_ctx, s, 32, vals, rows, cols, pkg.FP64)
    db = gpu_ctx.to_device(s.rhs)
    dx = gpu_ctx.empty(3 * s.n, np.float64)
    st = pkg.PCGSolver(gpu_ctx, tol=tol, stop_mode=stop_mode).solve(keep[1], keep[2], keep[3], rows.size, s.n, mas,
                                                                    db, dx)
    assert abs(st.iterations - itr) <= 1, (st.iterations, itr)
    if stop_mode == 1:
        assert st.rel_residual <= 1e-5
    mas.close()
    om.close()


# ------------------------------------------------------------- whole solve
@pytest.mark.parametrize("which", ["tiny", "small", "c0"])
def test_system_solve_matches_oracle(pkg, orc, which, scene_tiny, scene_small, scene_c0):
    s = {"tiny": scene_tiny, "small": scene_small, "c0": scene_c0}[which]
    ls = pkg.LinearSystem(s, bank_size=32)
    try:
        res = ls.solve(tol=1e-5)
        assert res.status == 0 and res.true_rel_residual <= 1e-5
        # the BCOO assembled on the device equals the oracle's
        gv, gr, gc = ls.sys.bcoo()
        ov, orr, oc = orc.assemble(s)
        np.testing.assert_array_equal(gr, orr)
        np.testing.assert_array_equal(gc, oc)
        assert rel(gv, ov) <= 1e-10
        # same PCG trajectory: iterate after a fixed number of iterations
        k = 15
        fixed = ls.solve(tol=1e-30, max_iter=k)
        xr = orc.solve_scene(s, 32, tol=1e-30, max_iter=k).x
        assert fixed.iterations == k - 1
        assert np.linalg.norm(fixed.x - xr) / np.linalg.norm(xr) <= 1e-8
        if which == "c0":
            # well-conditioned scene: tight solves agree to 1e-8 relative
            tight = ls.solve(tol=1e-11)
            ref = orc.solve_scene(s, 32, tol=1e-11)
            assert np.linalg.norm(tight.x - ref.x) / np.linalg.norm(ref.x) <= 1e-8
    finally:
        ls.close()


def test_mixed_precision_ir_converges(pkg, orc, scene_c0):
    s = scene_c0
    ls = pkg.LinearSystem(s, bank_size=32)
    try:
        mixed = ls.solve(tol=1e-5, precision="fp32_ir")
        plain = ls.solve(tol=1e-5, precision="fp64")
        assert mixed.true_rel_residual <= 1e-5, mixed
        ref = orc.solve_scene(s, 32, tol=1e-11)
        nrm = np.linalg.norm(ref.x)
        e_mixed = np.linalg.norm(mixed.x - ref.x) / nrm
        e_plain = np.linalg.norm(plain.x - ref.x) / nrm
        print(f"fp32_ir: iters={mixed.iterations} outer={mixed.outer_iterations} err={e_mixed:.3e}; "
              f"fp64: iters={plain.iterations} err={e_plain:.3e}")
        assert e_mixed <= 10 * e_plain + 1e-7
    finally:
        ls.close()


@pytest.mark.parametrize("fuse", [1, 2])
@pytest.mark.parametrize("precision", ["fp64", "fp32_mas"])
def test_fused_iteration_matches_unfused(pkg, scene_c0, precision, fuse):
    """The fused iteration (p update inside the SpMV, x/r update inside the
    first MAS kernel) performs the same arithmetic as the separate kernels;
    only the grouping of the ||r||^2 partials of the stop rule differs."""
    ls = pkg.LinearSystem(scene_c0, bank_size=32)
    try:
        k = 25
        a = ls.solve(tol=1e-30, max_iter=k, precision=precision, fuse=fuse)
        b = ls.solve(tol=1e-30, max_iter=k, precision=precision, fuse=0)
        assert a.iterations == b.iterations == k - 1
        assert np.linalg.norm(a.x - b.x) <= 1e-14 * np.linalg.norm(b.x)
        a = ls.solve(tol=1e-6, precision=precision, fuse=fuse)
        b = ls.solve(tol=1e-6, precision=precision, fuse=0)
        assert abs(a.iterations - b.iterations) <= 1, (a.iterations, b.iterations)
        assert a.true_rel_residual <= 1e-6 and b.true_rel_residual <= 1e-6
    finally:
        ls.close()


@pytest.mark.parametrize("coarse", [0, 1])
def test_flexible_pcg_matches_pcg_for_a_fixed_preconditioner(pkg, scene_c0, coarse):
    """With a fixed SPD preconditioner, beta = -(z.Ap)/(p.Ap) equals
    rz_new/rz_old in exact arithmetic: the flexible iteration follows PCG."""
    ls = pkg.LinearSystem(scene_c0, bank_size=32, coarse_mode=coarse)
    try:
        k = 20
        a = ls.solve(tol=1e-30, max_iter=k, flexible=True)
        b = ls.solve(tol=1e-30, max_iter=k, flexible=False)
        assert a.iterations == b.iterations == k - 1
        assert np.linalg.norm(a.x - b.x) <= 1e-8 * np.linalg.norm(b.x)
        a = ls.solve(tol=1e-6, flexible=True)
        b = ls.solve(tol=1e-6, flexible=False)
        assert abs(a.iterations - b.iterations) <= 1 and a.true_rel_residual <= 1e-6
    finally:
        ls.close()


@pytest.mark.parametrize("coarse", [0, 1])
def test_fp32_refinement_with_residual_replacement(pkg, scene_c0, coarse):
    """FP32-valued SpMV and FP32 MAS apply with the recursive residual
    replaced by the FP64 one every 16 iterations: one Krylov run meets the
    FP64 tolerance in about the iterations of the FP64 solve, fewer than
    refinement by restarts."""
    ls = pkg.LinearSystem(scene_c0, bank_size=32, coarse_mode=coarse)
    try:
        plain = ls.solve(tol=1e-6, precision="fp64")
        rep = ls.solve(tol=1e-6, precision="fp32_ir", ir_replace_every=16)
        rst = ls.solve(tol=1e-6, precision="fp32_ir", ir_replace_every=0)
        print(f"coarse={coarse}: fp64 {plain.iterations}, replacement {rep.iterations} "
              f"({rep.outer_iterations} pass), restarts {rst.iterations} ({rst.outer_iterations} passes)")
        assert rep.true_rel_residual <= 1e-6 and rst.true_rel_residual <= 1e-6
        assert rep.iterations <= rst.iterations
        assert rep.iterations <= 1.25 * plain.iterations + 8
        assert np.linalg.norm(rep.x - plain.x) <= 1e-4 * np.linalg.norm(plain.x)
    finally:
        ls.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
@pytest.mark.parametrize("coarse", [1, 2])
def test_graph_replay_matches_eager(pkg, precision, coarse):
    """PCG with its full chunks captured into a HIP graph and replayed gives
    bit-identical iterates to the eager loop (same kernels, same order; every
    kernel of the loop is deterministic).  Three graph solves, each followed
    by an eager one, so replays alternate with the program's own eager work."""
    s = pkg.scenes.tet_grid_scene("graph_grid16", 16, E_loguniform=(1e4, 1e9), n_pairs=800)
    ls = pkg.LinearSystem(s, bank_size=32, coarse_mode=coarse)
    try:
        ref = ls.solve(tol=1e-7, precision=precision, graph=False)
        assert ref.status == 0 and ref.iterations > 16  # several full chunks
        for _ in range(3):
            g = ls.solve(tol=1e-7, precision=precision, graph=True)
            assert g.status == 0 and g.iterations == ref.iterations
            assert np.array_equal(g.x, ref.x)
            e = ls.solve(tol=1e-7, precision=precision, graph=False)
            assert e.iterations == ref.iterations and np.array_equal(e.x, ref.x)
    finally:
        ls.close()


@pytest.mark.parametrize("form", ["halo", "gather"])
def test_spmv_long_rows_match_oracle(pkg, orc, gpu_ctx, scene_small, monkeypatch, form):
    """Rows of more than 128 blocks (here three hub rows joined to a few
    hundred vertices each, one of them the first row, one in the middle and
    one near the end) are kept outside the tiles and multiplied by their own
    kernel; every precision of the tiled product still equals the oracle's.
    With FP64 vectors and FP32 tile values the long rows still read their
    FP64 values (stiff hub rows, e.g. affine bodies, lose most in FP32), so
    those rows match the FP64 product."""
    s = scene_small
    vals, rows, cols = orc.assemble(s)
    rng = np.random.default_rng(5)
    extra_r, extra_c = [], []
    for hub, k in ((0, 300), (s.n // 2, 200), (s.n - 3, 150)):
        js = rng.choice(s.n, k, replace=False)
        for j in js:
            if j != hub:
                extra_r.append(min(hub, j))
                extra_c.append(max(hub, j))
    ev = rng.standard_normal((len(extra_r), 9)) * np.abs(vals).max() * 1e-3
    tv = np.concatenate([vals, ev])
    tr = np.concatenate([rows, np.array(extra_r, np.int32)])
    tc = np.concatenate([cols, np.array(extra_c, np.int32)])
    hv, hr, hc = orc.convert(tv, tr, tc)
    x = rng.standard_normal(3 * s.n)
    y0 = rng.standard_normal(3 * s.n)
    dv, dr, dc = upload_bcoo(gpu_ctx, hv, hr, hc)
    if form == "gather":
        monkeypatch.setenv("SGIPC_SPMV_GATHER", "1")
    A = pkg.SpMat(gpu_ctx, dv, dr, dc, hr.size, s.n, want_f32=True)
    monkeypatch.delenv("SGIPC_SPMV_GATHER", raising=False)
    try:
        assert A.long_rows() == 3
        assert A.layout()[0] == (1 if form == "halo" else 0)
        dx, dy = gpu_ctx.to_device(x), gpu_ctx.to_device(y0)
        A.apply(0.9, dx, -0.4, dy)
        ref = orc.spmv(0.9, hv, hr, hc, x, -0.4, y0, s.n)
        assert rel(dy.download(), ref) <= 1e-13
        x32 = gpu_ctx.to_device(x.astype(np.float32), np.float32)
        y32 = gpu_ctx.to_device(y0.astype(np.float32), np.float32)
        A.apply(0.9, x32, -0.4, y32)
        assert rel(y32.download().astype(np.float64), ref) <= 2e-6
        dy.upload(y0)
        A.apply_f32_values(0.9, dx, -0.4, dy)
        got = dy.download()
        assert rel(got, ref) <= 2e-6
        hubs = np.array([0, s.n // 2, s.n - 3])
        idx = (3 * hubs[:, None] + np.arange(3)).ravel()
        assert rel(got[idx], ref[idx]) <= 1e-13
    finally:
        A.close()
