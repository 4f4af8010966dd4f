"""Pins the oracle's affine-body (ABD) terms: the orthogonality energy
against finite differences and its rotation invariance, the body mass matrix
against sum m J^T J over the mass points, the contact class mapping against a
dense J^T H J of the point-level contact Hessian, the block-Jacobi inverses
against numpy, and the hybrid PCG against scipy's CG with the same
preconditioner."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from test_oracle_linalg import bcoo_to_csr


def J_of(xb):
    """3 x 12 ABD Jacobian of a point with rest coordinates xb."""
    J = np.zeros((3, 12))
    J[:, :3] = np.eye(3)
    for r in range(3):
        J[r, 3 + 3 * r:6 + 3 * r] = xb
    return J


def test_shape_energy_derivatives_and_invariance(orc):
    rng = np.random.default_rng(0)
    for _ in range(3):
        q = np.r_[rng.standard_normal(3), (np.eye(3) + 0.2 * rng.standard_normal((3, 3))).ravel()]
        e, g, H = orc.abd_shape(q)
        h = 1e-6
        fd = np.zeros(9)
        Hfd = np.zeros((9, 9))
        for k in range(9):
            qp, qm = q.copy(), q.copy()
            qp[3 + k] += h
            qm[3 + k] -= h
            fd[k] = (orc.abd_shape(qp)[0] - orc.abd_shape(qm)[0]) / (2 * h)
            Hfd[:, k] = (orc.abd_shape(qp)[1] - orc.abd_shape(qm)[1]) / (2 * h)
        assert np.max(np.abs(g - fd)) <= 1e-6 * np.max(np.abs(fd))
        assert np.max(np.abs(H - Hfd)) <= 1e-6 * np.max(np.abs(Hfd))
        assert np.max(np.abs(H - H.T)) <= 1e-12 * np.max(np.abs(H))
    # rotations are energy-free (closed form: the energy measures A A^T - I)
    th = 0.8
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    e, g, _ = orc.abd_shape(np.r_[np.zeros(3), R.ravel()])
    assert abs(e) <= 1e-28 and np.max(np.abs(g)) <= 1e-14
    A = np.diag([1.1, 0.9, 1.0])
    e, _, _ = orc.abd_shape(np.r_[np.zeros(3), A.ravel()])
    assert abs(e - ((1.21 - 1) ** 2 + (0.81 - 1) ** 2)) <= 1e-14


def test_body_hessian_mass_and_projection(orc, pkg):
    s = pkg.scenes.scene_hybrid_tiny()
    vals, rows, cols, grad = orc.abd_body(s)
    dt = s.dt
    for b in range(s.n_bodies):
        H = np.zeros((12, 12))
        for k in range(10):
            r, c = rows[10 * b + k] - 4 * b, cols[10 * b + k] - 4 * b
            blk = vals[10 * b + k].reshape(3, 3).T
            H[3 * r:3 * r + 3, 3 * c:3 * c + 3] = blk
            H[3 * c:3 * c + 3, 3 * r:3 * r + 3] = blk.T
        # mass matrix from the body's 8 corner mass points
        pts = s.x_bar[s.point_body == b]
        M = sum((s.mass10[b, 0] / 8) * J_of(xb).T @ J_of(xb) for xb in pts)
        if s.body_fixed[b]:
            assert np.max(np.abs(H - M)) <= 1e-12 * np.abs(M).max()
            assert np.all(grad[12 * b:12 * b + 12] == 0)
            continue
        kvt2 = s.kappa_abd * s.body_volume[b] * dt * dt
        e, g, Hs = orc.abd_shape(s.q[b])
        w, V = np.linalg.eigh(kvt2 * Hs)
        P = (V * np.maximum(w, 0)) @ V.T
        ref = M.copy()
        ref[3:, 3:] += P
        assert np.max(np.abs(H - ref)) <= 1e-9 * np.abs(ref).max()
        gref = M @ (s.q[b] - s.q_tilde[b])
        gref[3:] += kvt2 * g
        assert np.max(np.abs(grad[12 * b:12 * b + 12] - gref)) <= 1e-9 * np.abs(gref).max()


def _point_jacobian(s):
    """(3 * n_points) x (3 * n_sys) map from system DoFs to point positions."""
    n_pts = s.n_abd_points + s.n
    Jg = np.zeros((3 * n_pts, 3 * s.n_sys))
    for i in range(s.n_abd_points):
        b = s.point_body[i]
        Jg[3 * i:3 * i + 3, 12 * b:12 * b + 12] = 0 if s.body_fixed[b] else J_of(s.x_bar[i])
    for v in range(s.n):
        Jg[3 * (s.n_abd_points + v):3 * (s.n_abd_points + v) + 3,
           3 * (4 * s.n_bodies + v):3 * (4 * s.n_bodies + v) + 3] = np.eye(3)
    return Jg


def test_contact_map_is_jt_h_j(orc, pkg):
    s = pkg.scenes.scene_hybrid_tiny(N=5, n_bodies=4, n_pairs=0)
    rng = np.random.default_rng(5)
    # hand-made pairs: body point vs FEM triangle, body point vs another body's
    # points, body point vs its own body's points, FEM point vs body points
    P = s.n_abd_points
    pairs = [[-(0) - 1, P + 0, P + 1, P + 2], [-(9) - 1, 0, 1, 2], [-(3) - 1, 0, 1, 2],
             [-(P + 5) - 1, 8, 9, 10], [-(17) - 1, P + 3, P + 4, 25], [-(30) - 1, P + 7, 9, 31]]
    s.pairs = np.array(pairs, np.int32)
    s.body_fixed[:] = 0
    s.body_fixed[3] = 1
    geo = rng.standard_normal((len(pairs), 4, 3)) * 1e-3
    geo[:, 0] = geo[:, 1:].mean(axis=1) + np.array([0, 0, 5e-4])
    s.pair_x = geo
    bv, br, bc, _ = orc.barrier_pt(s, with_grad=False)
    n_pts = P + s.n
    Hp = bcoo_to_csr(bv, br, bc, n_pts).toarray()     # symmetric point-level contact Hessian
    Jg = _point_jacobian(s)
    ref = Jg.T @ Hp @ Jg
    mv, mr, mc = orc.abd_contact_map(s, bv, br, bc)
    assert np.all(mr <= mc)
    ov, orr, oc = orc.convert(mv, mr, mc)
    got = bcoo_to_csr(ov, orr, oc, s.n_sys).toarray()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(ref).max()


def test_block_jacobi_and_hybrid_pcg(orc, pkg):
    s = pkg.scenes.scene_hybrid_tiny()
    vals, rows, cols = orc.assemble(s)
    A = bcoo_to_csr(vals, rows, cols, s.n_sys)
    nb = s.n_bodies
    inv = orc.abd_jacobi(vals, rows, cols, nb)
    Ad = A.toarray()
    for b in range(nb):
        blk = Ad[12 * b:12 * b + 12, 12 * b:12 * b + 12]
        assert np.max(np.abs(inv[b] - np.linalg.inv(blk))) <= 1e-9 * np.abs(inv[b]).max()
    m = orc.Mas(s, 32)
    m.build(vals, rows, cols, 4 * nb, s.pairs)

    def Mv(v):
        v = np.asarray(v).reshape(-1)
        z = v.copy()
        for b in range(nb):
            z[12 * b:12 * b + 12] = inv[b] @ v[12 * b:12 * b + 12]
        z[12 * nb:] = m.apply(v[12 * nb:])
        return z
    Mop = spla.LinearOperator(A.shape, matvec=Mv)
    for k in (3, 12):
        xr, it, _ = orc.pcg2(vals, rows, cols, s.n_sys, s.rhs, m, 4 * nb, inv, tol=1e-30, max_iter=k + 1)
        xs, _ = spla.cg(A, s.rhs, x0=np.zeros_like(s.rhs), rtol=1e-30, atol=0.0, maxiter=k, M=Mop)
        assert it == k
        assert np.linalg.norm(xr - xs) <= 1e-8 * np.linalg.norm(xs), k
    res = orc.solve_scene(s, 32, tol=1e-5)
    assert res.true_rel_residual <= 1e-5
    assert np.linalg.norm(s.rhs - A @ res.x) / np.linalg.norm(s.rhs) <= 1e-5 + 1e-12
    m.close()


def _one_body_scene(pkg, q, q_tilde, fixed, theta_speed=31.4, strength=10.0, dt=0.01):
    """A hybrid scene reduced to its body terms: one body with the given q."""
    s = pkg.scenes.scene_hybrid_tiny(N=4, n_bodies=1, n_pairs=0)
    import dataclasses
    return dataclasses.replace(s, q=q[None, :], q_tilde=q_tilde[None, :], body_fixed=np.array([fixed], np.int32),
                               motor_speed=theta_speed, motor_strength=strength, dt=dt)


def test_motor_term_closed_form(orc, pkg):
    """Motor body (setup_abd_system_gradient_and_hessian.cu:271-339) with
    A~ = I: the target rows a2, a3 of q_p are the rows of R(theta) (the rest
    frame turned about x), so the motor adds W (q - q_R) to the a2, a3 rows
    of G and W to H there, W = strength * M[6:12, 6:12]; the motor body
    differs from the same body set free by exactly that."""
    rng = np.random.default_rng(3)
    for speed in (31.4, 100.0):
        qt = np.r_[rng.standard_normal(3), np.eye(3).ravel()]
        q = qt + 0.01 * rng.standard_normal(12)
        dt, strength = 0.01, 7.0
        mot = _one_body_scene(pkg, q, qt, 2, speed, strength, dt)
        free = _one_body_scene(pkg, q, qt, 0, speed, strength, dt)
        vm, _, _, gm = orc.abd_body(mot)
        vf, _, _, gf = orc.abd_body(free)
        th = speed * dt
        c, s_ = np.cos(th), np.sin(th)
        target = np.array([[0, c, -s_], [0, s_, c]]).ravel()  # rows 1, 2 of R
        M = np.zeros((12, 12))
        m10 = mot.mass10[0]
        M[:3, :3] = m10[0] * np.eye(3)
        D = np.array([[m10[4], m10[5], m10[6]], [m10[5], m10[7], m10[8]], [m10[6], m10[8], m10[9]]])
        for r in range(3):
            M[r, 3 + 3 * r:6 + 3 * r] = m10[1:4]
            M[3 + 3 * r:6 + 3 * r, r] = m10[1:4]
            M[3 + 3 * r:6 + 3 * r, 3 + 3 * r:6 + 3 * r] = D
        W = strength * M[6:, 6:]
        expect = np.zeros(12)
        expect[6:] = W @ (q[6:] - target)
        assert np.max(np.abs((gm - gf) - expect)) <= 1e-12 * np.max(np.abs(expect))
        # the Hessian difference is W on the (a2, a3) blocks: blocks (2,2), (2,3), (3,3)
        dv = (vm - vf).reshape(10, 3, 3)          # upper blocks of rows 0..3, column-major
        blocks = {(i, j): k for k, (i, j) in enumerate((i, j) for i in range(4) for j in range(i, 4))}
        for (i, j), k in blocks.items():
            want = W[3 * (i - 2):3 * (i - 1), 3 * (j - 2):3 * (j - 1)] if i >= 2 and j >= 2 else np.zeros((3, 3))
            # (H + W) - H rounds at the size of H's shape block
            assert np.max(np.abs(dv[k].T - want)) <= 1e-12 * np.max(np.abs(vm)), (i, j)


def test_dirichlet_assembly_projects_the_pinned_vertices(orc, pkg):
    """With pinned vertices the assembled matrix equals the free one on the
    free rows and columns, and holds only the mass block on a pinned vertex's
    rows and columns (GIPC.cu:10332-10363,10444-10462); the gradient is the
    free one with the pinned rows zeroed (fem_linear_subsystem.cu:73-76).
    Checked on dense matrices of a small contact scene and of a hybrid one."""
    for s0 in (pkg.scenes.tet_grid_scene("bc6", 6, E_loguniform=(1e4, 1e9), n_pairs=30),
               pkg.scenes.scene_hybrid_tiny(N=6, n_bodies=4, n_pairs=30)):
        s1 = pkg.scenes.with_dirichlet(s0)
        n_sys = getattr(s0, "n_sys", s0.n)
        off = 4 * getattr(s0, "n_bodies", 0)
        H0 = bcoo_to_csr(*orc.assemble(s0), n_sys).toarray()
        H1 = bcoo_to_csr(*orc.assemble(s1), n_sys).toarray()
        _, _, _, g0 = orc.triplets(s0, with_grad=True)
        _, _, _, g1 = orc.triplets(s1, with_grad=True)
        pin = np.zeros(n_sys, bool)
        pin[off:] = s1.boundary_type != 0
        pin3 = np.repeat(pin, 3)
        free = ~pin3
        assert pin.sum() > 0
        assert np.max(np.abs(H1[np.ix_(free, free)] - H0[np.ix_(free, free)])) <= 1e-12 * np.abs(H0).max()
        assert np.all(H1[np.ix_(pin3, free)] == 0) and np.all(H1[np.ix_(free, pin3)] == 0)
        m = np.repeat(s0.mass[pin[off:]], 3)
        np.testing.assert_array_equal(H1[np.ix_(pin3, pin3)], np.diag(m))
        np.testing.assert_array_equal(g1[pin3], 0.0)
        assert np.max(np.abs(g1[free] - g0[free])) <= 1e-13 * np.abs(g0).max()


def test_hybrid_gradient_is_jt_of_the_point_gradient(orc, pkg):
    """The body rows of the assembled gradient minus the body term equal
    sum J(xbar_i)^T g_i over the body's points (dense 3 x 12 Jacobians), for
    non-fixed bodies; fixed bodies carry zero."""
    s = pkg.scenes.scene_hybrid_tiny(N=6, n_bodies=5, n_pairs=60)
    rng = np.random.default_rng(2)
    gp = rng.standard_normal(3 * (s.n_abd_points + s.n))
    g = orc.point_gradient_to_system(s, gp)
    want = np.zeros(12 * s.n_bodies)
    for i in range(s.n_abd_points):
        b = s.point_body[i]
        if s.body_fixed[b] == 1:
            continue
        want[12 * b:12 * b + 12] += J_of(s.x_bar[i]).T @ gp[3 * i:3 * i + 3]
    assert np.max(np.abs(g[:12 * s.n_bodies] - want)) <= 1e-13 * np.abs(want).max()
    np.testing.assert_array_equal(g[12 * s.n_bodies:], gp[3 * s.n_abd_points:])
