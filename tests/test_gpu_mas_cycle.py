"""MAS with the level-1 V(1,1) cycle (coarse mode 1) against its numpy
restatement (oracle/mpnn.py: mas_cycle_apply), symmetry, and solves."""
import numpy as np
import pytest

from oracle import mpnn
from test_gpu_parity import _gpu_mas

pytestmark = pytest.mark.gpu


def _cycle_mas(pkg, ctx, s, B, vals, rows, cols, precision):
    """The device MAS with the cycle switched on before the build."""
    nbr = [ctx.to_device(a, np.uint32) for a in (s.nbr_start, s.nbr_num,
                                                 s.nbr_list if s.nbr_list.size else np.zeros(1, np.uint32))]
    mas = pkg.MASPreconditioner(ctx, B)
    mas.set_coarse_mode(1)
    mas.initPreconditioner_Neighbor(s.n, 0, int(s.nbr_list.size), *nbr)
    mas.initPreconditioner_Matrix()
    dv = ctx.to_device(vals, np.float64)
    dr = ctx.to_device(rows, np.int32)
    dc = ctx.to_device(cols, np.int32)
    pairs = ctx.to_device(s.pairs.reshape(-1), np.int32) if s.n_pairs else None
    mas.setPreconditioner_bcoo(dv, dr, dc, rows.size, 0, pairs, s.n_pairs, precision)
    return mas, (nbr, dv, dr, dc, pairs)


@pytest.mark.parametrize("B", [16, 32])
@pytest.mark.parametrize("which", ["tiny", "small", "c0"])
def test_cycle_apply_matches_oracle(pkg, orc, gpu_ctx, B, which, scene_tiny, scene_small, scene_c0):
    s = {"tiny": scene_tiny, "small": scene_small, "c0": scene_c0}[which]
    vals, rows, cols = orc.assemble(s)
    om = orc.Mas(s, B)
    om.build(vals, rows, cols, 0, s.pairs)
    parts = mpnn.MasParts(om)
    if parts.L <= 2 or parts.cnt[1] <= B:
        pytest.skip("level 1 is one domain (solved exactly): the cycle stays off")
    mas, keep = _cycle_mas(pkg, gpu_ctx, s, B, vals, rows, cols, pkg.FP64)
    try:
        # the restatement with the device's own domain inverses: what remains
        # is summation order and the float32 level-1 operator (in both)
        parts.inv = mpnn._domain_dense(mas.inverse_blocks().reshape(-1, B * (B + 1) // 2, 9), B)
        A = mpnn.bcoo_to_csr(vals, rows, cols, s.n)
        graph = mpnn.coarse_graph(A, parts.map1, parts.cnt[1])
        A1f = mpnn.float_operator(graph)
        A2 = mpnn.level2_operator(parts, graph)
        rng = np.random.default_rng(4)
        for k in range(3):
            r = rng.standard_normal(3 * s.n)
            ref = mpnn.mas_cycle_apply(parts, A1f, r, A2=A2)
            dr_ = gpu_ctx.to_device(r)
            dz = gpu_ctx.empty(3 * s.n, np.float64)
            mas.preconditioning(dr_, dz)
            got = dz.download()
            err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            # float32 level-1 operator on both sides; what remains is the
            # summation order and the Cholesky vs LU level-2 solve
            assert err <= 1e-7, (k, err)
    finally:
        mas.close()
        om.close()


def level2_inverse_check(pkg, orc, mas, om, s, vals, rows, cols):
    """The device's dense level-2 inverse against numpy.linalg.inv of the
    oracle's FP64 Galerkin operator P_12^T A_1 P_12: its residual
    max|A_2 G - I| within 10x numpy's own, the entries within 1e-6 of the
    largest, symmetric to 1e-9."""
    parts = mpnn.MasParts(om)
    A = mpnn.bcoo_to_csr(vals, rows, cols, s.n)
    graph = mpnn.coarse_graph(A, parts.map1, parts.cnt[1])
    A2 = mpnn.level2_operator(parts, graph)
    G = mas.level2_inverse()
    assert G is not None and G.shape == A2.shape, (None if G is None else G.shape, A2.shape)
    ref = np.linalg.inv(A2)
    eye = np.eye(A2.shape[0])
    res_gpu = np.abs(A2 @ G - eye).max()
    res_np = np.abs(A2 @ ref - eye).max()
    diff = np.abs(G - ref).max() / np.abs(ref).max()
    asym = np.abs(G - G.T).max() / np.abs(G).max()
    print(f"level 2: m={A2.shape[0]} residual {res_gpu:.2e} (numpy {res_np:.2e}) diff {diff:.2e} asym {asym:.2e}")
    assert res_gpu <= 10 * res_np + 1e-12
    assert diff <= 1e-6 and asym <= 1e-9
    return A2.shape[0]


@pytest.mark.parametrize("N", [24, 32])
def test_level2_inverse_matches_numpy(pkg, orc, gpu_ctx, N):
    """Blocked Gauss-Jordan over 96 x 96 tiles: 24^3 at B = 16 gives a level 2
    of 2 tiles with padding, 32^3 one of 4 tiles."""
    s = pkg.scenes.tet_grid_scene(f"l2_grid{N}", N, E_loguniform=(1e4, 1e9), n_pairs=20 * N * N)
    B = 16
    vals, rows, cols = orc.assemble(s)
    om = orc.Mas(s, B)
    om.build(vals, rows, cols, 0, s.pairs)
    mas, keep = _cycle_mas(pkg, gpu_ctx, s, B, vals, rows, cols, pkg.FP64)
    try:
        m = level2_inverse_check(pkg, orc, mas, om, s, vals, rows, cols)
        assert m > 96
    finally:
        mas.close()
        om.close()


def test_cycle_operator_is_spd(pkg, orc, gpu_ctx):
    s = pkg.scenes.tet_grid_scene("small_grid8", 8, E_loguniform=(1e4, 1e9), n_pairs=40)
    B = 16
    vals, rows, cols = orc.assemble(s)
    mas, keep = _cycle_mas(pkg, gpu_ctx, s, B, vals, rows, cols, pkg.FP64)
    try:
        n3 = 3 * s.n
        M = np.zeros((n3, n3))
        dz = gpu_ctx.empty(n3, np.float64)
        for j in range(n3):
            e = np.zeros(n3)
            e[j] = 1.0
            mas.preconditioning(gpu_ctx.to_device(e), dz)
            M[:, j] = dz.download()
        asym = np.abs(M - M.T).max() / np.abs(M).max()
        assert asym <= 1e-9, asym
        assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0
    finally:
        mas.close()


@pytest.mark.parametrize("B", [16, 32])
def test_cycle_solves_in_fewer_iterations(pkg, B):
    s = pkg.scenes.tet_grid_scene("cycle_grid16", 16, E_loguniform=(1e4, 1e9), n_pairs=800)
    plain = pkg.LinearSystem(s, bank_size=B, coarse_mode=0)
    cyc = pkg.LinearSystem(s, bank_size=B, coarse_mode=1)
    try:
        a = plain.solve(tol=1e-5)
        b = cyc.solve(tol=1e-5)
        print(f"B={B}: additive {a.iterations} iterations, cycle {b.iterations}")
        assert a.status == 0 and b.status == 0
        assert b.true_rel_residual <= 1e-5
        assert b.iterations < a.iterations
        # same solution to the solve tolerance
        assert np.linalg.norm(a.x - b.x) <= 1e-4 * np.linalg.norm(a.x)
    finally:
        plain.close()
        cyc.close()


# ---- coarse mode 2: two-grid cycle on level 0 around the level-1 cycle
def _vcycle_mas(pkg, ctx, s, B, vals, rows, cols, precision, omega=None, half=None):
    nbr = [ctx.to_device(a, np.uint32) for a in (s.nbr_start, s.nbr_num,
                                                 s.nbr_list if s.nbr_list.size else np.zeros(1, np.uint32))]
    mas = pkg.MASPreconditioner(ctx, B)
    mas.set_coarse_mode(2)
    if omega is not None:
        mas.set_smoothing_weight(omega)
    if half is not None:
        mas.set_level0_half(half)
    mas.initPreconditioner_Neighbor(s.n, 0, int(s.nbr_list.size), *nbr)
    mas.initPreconditioner_Matrix()
    dv = ctx.to_device(vals, np.float64)
    dr = ctx.to_device(rows, np.int32)
    dc = ctx.to_device(cols, np.int32)
    pairs = ctx.to_device(s.pairs.reshape(-1), np.int32) if s.n_pairs else None
    mas.setPreconditioner_bcoo(dv, dr, dc, rows.size, 0, pairs, s.n_pairs, precision)
    return mas, (nbr, dv, dr, dc, pairs)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32_half"])
@pytest.mark.parametrize("B", [16, 32])
@pytest.mark.parametrize("which", ["small", "c0"])
def test_vcycle_apply_matches_oracle(pkg, orc, gpu_ctx, B, which, precision, scene_small, scene_c0):
    """The device apply of coarse mode 2 against mpnn.mas_vcycle_apply with the
    device's own domain inverses.  FP64: what remains is summation order and
    the float32 level-1 operator and A P (rounded the same way in both).
    FP32 (slots, inverses, the fine SpMV's values and vectors in float32):
    within 2e-4 of the FP64 restatement.  FP32 with the FP16 level-0 storage:
    within 2e-4 of the restatement given the same FP16 level-0 matrices, and
    within 3e-3 of the one with the float inverses (the storage's own error:
    FP16 keeps 11 significant bits of each scaled entry)."""
    s = {"small": scene_small, "c0": scene_c0}[which]
    vals, rows, cols = orc.assemble(s)
    om = orc.Mas(s, B)
    om.build(vals, rows, cols, 0, s.pairs)
    parts = mpnn.MasParts(om)
    if parts.L <= 2 or parts.cnt[1] <= B:
        pytest.skip("level 1 is one domain (solved exactly): the cycle stays off")
    prec = pkg.FP64 if precision == "fp64" else pkg.FP32_MAS
    half = precision == "fp32_half"
    mas, keep = _vcycle_mas(pkg, gpu_ctx, s, B, vals, rows, cols, prec, omega=0.45, half=half)
    try:
        parts.inv = mpnn._domain_dense(mas.inverse_blocks().reshape(-1, B * (B + 1) // 2, 9), B)
        inv_float = parts.inv
        if half:
            parts.inv = mpnn.half_level0_inverse(inv_float, parts.off[1] // B)
        A = mpnn.bcoo_to_csr(vals, rows, cols, s.n)
        graph = mpnn.coarse_graph(A, parts.map1, parts.cnt[1])
        A1f = mpnn.float_operator(graph)
        A2 = mpnn.level2_operator(parts, graph)
        APf = mpnn.ap_operator(A, parts.map1, parts.cnt[1])
        rng = np.random.default_rng(5)
        for k in range(3):
            r = rng.standard_normal(3 * s.n)
            ref = mpnn.mas_vcycle_apply(parts, A1f, A, APf, r, A2=A2, omega=0.45)
            dz = gpu_ctx.empty(3 * s.n, np.float64)
            mas.preconditioning(gpu_ctx.to_device(r), dz)
            got = dz.download()
            err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            assert err <= (1e-7 if precision == "fp64" else 2e-4), (k, err)
            if half:
                parts.inv = inv_float
                ref_f = mpnn.mas_vcycle_apply(parts, A1f, A, APf, r, A2=A2, omega=0.45)
                parts.inv = mpnn.half_level0_inverse(inv_float, parts.off[1] // B)
                err_f = np.linalg.norm(got - ref_f) / np.linalg.norm(ref_f)
                print(f"fp16 level 0: {err:.2e} vs its restatement, {err_f:.2e} vs float storage")
                assert err_f <= 3e-3, (k, err_f)
    finally:
        mas.close()
        om.close()


def test_vcycle_operator_is_spd(pkg, orc, gpu_ctx):
    s = pkg.scenes.tet_grid_scene("small_grid8", 8, E_loguniform=(1e4, 1e9), n_pairs=40)
    B = 16
    vals, rows, cols = orc.assemble(s)
    mas, keep = _vcycle_mas(pkg, gpu_ctx, s, B, vals, rows, cols, pkg.FP64)
    try:
        n3 = 3 * s.n
        M = np.zeros((n3, n3))
        dz = gpu_ctx.empty(n3, np.float64)
        for j in range(n3):
            e = np.zeros(n3)
            e[j] = 1.0
            mas.preconditioning(gpu_ctx.to_device(e), dz)
            M[:, j] = dz.download()
        # A P in float32 makes the two residual paths differ by ~1e-7
        asym = np.abs(M - M.T).max() / np.abs(M).max()
        assert asym <= 1e-6, asym
        assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0
    finally:
        mas.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
@pytest.mark.parametrize("B", [16, 32])
def test_vcycle_solves_in_fewer_iterations(pkg, B, precision):
    s = pkg.scenes.tet_grid_scene("cycle_grid16", 16, E_loguniform=(1e4, 1e9), n_pairs=800)
    cyc = pkg.LinearSystem(s, bank_size=B, coarse_mode=1)
    two = pkg.LinearSystem(s, bank_size=B, coarse_mode=2)
    try:
        a = cyc.solve(tol=1e-5, precision=precision)
        b = two.solve(tol=1e-5, precision=precision)
        print(f"B={B} {precision}: level-1 cycle {a.iterations} iterations, two-grid on level 0 {b.iterations}")
        assert a.status == 0 and b.status == 0
        assert b.true_rel_residual <= 1e-5
        assert b.iterations < a.iterations
        assert np.linalg.norm(a.x - b.x) <= 1e-4 * np.linalg.norm(a.x)
        # the solve statistics name the path: mode 2 with the exact level-2 solve
        assert (b.coarse_mode, b.level2_exact, b.omega) == (2, True, 0.4)
        assert (a.coarse_mode, a.level2_exact) == (1, True)
    finally:
        cyc.close()
        two.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_ir"])
def test_vcycle_recovers_from_a_too_large_weight(pkg, precision):
    """With w = 1.9 the two-grid preconditioner is indefinite (w exceeds
    2 / lambda_max(M0 A)): the solver must still return a converged solution
    (PCG stops on r.z <= 0 and the system retries with half the weight)."""
    s = pkg.scenes.tet_grid_scene("cycle_grid16", 16, E_loguniform=(1e4, 1e9), n_pairs=800)
    ls = pkg.LinearSystem(s, bank_size=16, coarse_mode=2)
    ref = pkg.LinearSystem(s, bank_size=16, coarse_mode=2)
    try:
        ls.set_smoothing_weight(1.9)
        a = ls.solve(tol=1e-5, precision=precision)
        b = ref.solve(tol=1e-5, precision=precision)
        print(f"{precision}: w=1.9 start -> {a.iterations} iterations (mode {a.coarse_mode}, w {a.omega}); "
              f"default w: {b.iterations}")
        assert a.status == 0 and a.true_rel_residual <= 1e-5
        # the retries halved the weight (kept), or fell back to mode 1
        assert a.omega < 1.9 and a.coarse_mode in (1, 2)
        assert np.linalg.norm(a.x - b.x) <= 1e-4 * np.linalg.norm(b.x)
    finally:
        ls.close()
        ref.close()


def test_vcycle_weight_after_a_recovery_stays_put(pkg):
    """Two solves in a row from w = 1.9: a weight that the halving cured is
    kept, and the second solve then runs the two-grid cycle once, in the
    iterations of a fresh system given that weight directly; when only the
    level-1 fallback cured it, the weight is restored, not left halved."""
    s = pkg.scenes.tet_grid_scene("cycle_grid16", 16, E_loguniform=(1e4, 1e9), n_pairs=800)
    ls = pkg.LinearSystem(s, bank_size=16, coarse_mode=2)
    try:
        ls.set_smoothing_weight(1.9)
        a = ls.solve(tol=1e-5, precision="fp32_ir")
        b = ls.solve(tol=1e-5, precision="fp32_ir")
        print(f"first: mode {a.coarse_mode} w {a.omega} {a.iterations} it; second: mode {b.coarse_mode} "
              f"w {b.omega} {b.iterations} it")
        assert a.status == 0 and b.status == 0 and b.true_rel_residual <= 1e-5
        assert b.omega == a.omega
        if a.coarse_mode == 2:
            assert 0.05 <= a.omega < 1.9 and b.coarse_mode == 2
            fresh = pkg.LinearSystem(s, bank_size=16, coarse_mode=2)
            try:
                fresh.set_smoothing_weight(a.omega)
                c = fresh.solve(tol=1e-5, precision="fp32_ir")
            finally:
                fresh.close()
            assert b.iterations == c.iterations
        else:
            assert a.coarse_mode == 1 and a.omega == 1.9
    finally:
        ls.close()


@pytest.mark.parametrize("B", [16, 32])
def test_level0_half_storage_keeps_the_iteration_count(pkg, B):
    """FP16 level-0 inverses (the default of the float applies of coarse mode
    2) against float ones on a stiff contact scene: both converge to the
    requested true residual, within 3 iterations of each other, to the same
    solution."""
    s = pkg.scenes.tet_grid_scene("cycle_grid16", 16, E_loguniform=(1e4, 1e9), n_pairs=800)
    a_ls = pkg.LinearSystem(s, bank_size=B, coarse_mode=2)
    b_ls = pkg.LinearSystem(s, bank_size=B, coarse_mode=2)
    try:
        a_ls.set_level0_half(False)
        b_ls.set_level0_half(True)
        a = a_ls.solve(tol=1e-5, precision="fp32_ir")
        b = b_ls.solve(tol=1e-5, precision="fp32_ir")
        print(f"B={B}: float level 0 {a.iterations} iterations, fp16 level 0 {b.iterations}")
        assert a.status == 0 and b.status == 0
        assert b.true_rel_residual <= 1e-5
        assert b.iterations <= a.iterations + 3
        assert np.linalg.norm(a.x - b.x) <= 1e-4 * np.linalg.norm(a.x)
    finally:
        a_ls.close()
        b_ls.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_vcycle_level1_sums_in_the_spmv_epilogue(pkg, orc, gpu_ctx, monkeypatch, precision):
    """Coarse mode 2 forms the level-1 sums of r1 in the epilogue of the halo
    SpMV (r1 = r - w A z0); with the gather form of the fine matrix they come
    from their own kernel.  Both applies agree to the SpMV forms' summation
    order (FP64 1e-12, FP32 1e-5)."""
    s = pkg.scenes.tet_grid_scene("cycle_grid16", 16, E_loguniform=(1e4, 1e9), n_pairs=800)
    B = 32
    vals, rows, cols = orc.assemble(s)
    prec = pkg.FP64 if precision == "fp64" else pkg.FP32_MAS
    out = []
    for gather in (False, True):
        if gather:
            monkeypatch.setenv("SGIPC_SPMV_GATHER", "1")
        else:
            monkeypatch.delenv("SGIPC_SPMV_GATHER", raising=False)
        mas, keep = _vcycle_mas(pkg, gpu_ctx, s, B, vals, rows, cols, prec)
        try:
            rng = np.random.default_rng(11)
            zs = []
            for k in range(2):
                dz = gpu_ctx.empty(3 * s.n, np.float64)
                mas.preconditioning(gpu_ctx.to_device(rng.standard_normal(3 * s.n)), dz)
                zs.append(dz.download())
            out.append(zs)
        finally:
            mas.close()
    monkeypatch.delenv("SGIPC_SPMV_GATHER", raising=False)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for a, b in zip(*out):
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"{precision}: fused vs separate level-1 sums {err:.2e}")
        assert err <= tol, err


@pytest.mark.parametrize("which", ["c0", "grid20_contact"])
def test_vcycle_pcg_matches_restated_pcg(pkg, orc, gpu_ctx, which, scene_c0):
    """The headline preconditioner inside the whole PCG: the device's FP64
    PCG with coarse mode 2 against the reference's PCG loop (cemas.pcg, the
    loop of pcg_solver.cu:191-332) around the numpy restatement of the
    two-grid cycle, given the device's own domain inverses.  After a fixed
    11 iterations the iterates agree to 1e-6 (the applies agree to 1e-7);
    run to ||r||/||b|| <= 1e-5 the iteration counts are at most one apart and
    both solutions meet the tolerance."""
    from oracle import cemas
    if which == "c0":
        s = scene_c0
    else:
        s = pkg.scenes.tet_grid_scene("cycle_pcg20", 20, E_loguniform=(1e4, 1e9), n_pairs=1500)
    B = 32
    vals, rows, cols = orc.assemble(s)
    om = orc.Mas(s, B)
    om.build(vals, rows, cols, 0, s.pairs)
    parts = mpnn.MasParts(om)
    if parts.L <= 2 or parts.cnt[1] <= B:
        pytest.skip("level 1 is one domain (solved exactly): the cycle stays off")
    mas, keep = _vcycle_mas(pkg, gpu_ctx, s, B, vals, rows, cols, pkg.FP64, omega=0.4)
    try:
        parts.inv = mpnn._domain_dense(mas.inverse_blocks().reshape(-1, B * (B + 1) // 2, 9), B)
        A = mpnn.bcoo_to_csr(vals, rows, cols, s.n)
        graph = mpnn.coarse_graph(A, parts.map1, parts.cnt[1])
        A1f = mpnn.float_operator(graph)
        A2 = mpnn.level2_operator(parts, graph)
        APf = mpnn.ap_operator(A, parts.map1, parts.cnt[1])
        M = lambda r: mpnn.mas_vcycle_apply(parts, A1f, A, APf, r, A2=A2, omega=0.4)
        db = gpu_ctx.to_device(s.rhs)
        # fixed iteration count
        k = 12
        xr, itr = cemas.pcg(A, s.rhs, M, tol=1e-30, stop_mode=1, max_iter=k)
        dx = gpu_ctx.empty(3 * s.n, np.float64)
        st = pkg.PCGSolver(gpu_ctx, tol=1e-30, max_iter=k).solve(keep[1], keep[2], keep[3], rows.size, s.n, mas,
                                                                db, dx)
        assert st.iterations == itr == k - 1
        err = np.linalg.norm(dx.download() - xr) / np.linalg.norm(xr)
        assert err <= 1e-6, err
        # to the tolerance
        xr, itr = cemas.pcg(A, s.rhs, M, tol=1e-5, stop_mode=1)
        dx = gpu_ctx.empty(3 * s.n, np.float64)
        st = pkg.PCGSolver(gpu_ctx, tol=1e-5, stop_mode=1).solve(keep[1], keep[2], keep[3], rows.size, s.n, mas,
                                                                 db, dx)
        print(f"{which}: device {st.iterations} iterations, restated {itr}")
        assert abs(st.iterations - itr) <= 1, (st.iterations, itr)
        assert st.rel_residual <= 1e-5
        assert np.linalg.norm(s.rhs - A @ xr) / np.linalg.norm(s.rhs) <= 1e-5
    finally:
        mas.close()
        om.close()
