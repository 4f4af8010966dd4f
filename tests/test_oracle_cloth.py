"""Pins the oracle's cloth terms (Baraff-Witkin stretch/shear triangles and
quadratic bending) against finite differences, numpy eigen-projections and
closed-form special cases."""
import numpy as np
import pytest


def F_to_x9(F):
    """Triangle with x0 = 0, x1 = f_u, x2 = f_v: F = Ds when Dm^-1 = I."""
    return np.r_[np.zeros(3), F[:, 0], F[:, 1]]


def energy_F(orc, F, kst, ksh, rate):
    return orc.tri_bw_energy(F_to_x9(F), np.eye(2).ravel(), 1.0, kst, ksh, rate)


def eig_clamp(H):
    H = 0.5 * (H + H.T)
    w, V = np.linalg.eigh(H)
    return (V * np.maximum(w, 0.0)) @ V.T


def fd_hessian_F(orc, F, kst, ksh, rate, h=1e-6):
    """Central differences of dE/dF (6 x 6, column-major vec(F))."""
    H = np.zeros((6, 6))
    f = F.T.ravel().copy()  # column-major vec
    for q in range(6):
        fp, fm = f.copy(), f.copy()
        fp[q] += h
        fm[q] -= h
        Pp, _ = orc.bw_dpdf(fp, kst, ksh, rate)
        Pm, _ = orc.bw_dpdf(fm, kst, ksh, rate)
        H[:, q] = (Pp - Pm) / (2 * h)
    return H


SAMPLES = [
    np.array([[1.3, 0.2], [0.1, 0.9], [0.2, -0.1]]),   # stretch along u, compress v, sheared
    np.array([[0.8, -0.3], [0.2, 0.7], [-0.1, 0.4]]),  # compressed
    np.array([[1.05, 0.01], [0.0, 1.2], [0.3, 0.1]]),
]


@pytest.mark.parametrize("k", range(len(SAMPLES)))
def test_bw_stress_is_energy_gradient(orc, k):
    F = SAMPLES[k]
    kst, ksh, rate = 3.3e5, 1.0e6, 100.0
    P, _ = orc.bw_dpdf(F.T.ravel(), kst, ksh, rate)
    h = 1e-7
    fd = np.zeros(6)
    for q in range(6):
        Fp, Fm = F.copy(), F.copy()
        Fp[q % 3, q // 3] += h
        Fm[q % 3, q // 3] -= h
        fd[q] = (energy_F(orc, Fp, kst, ksh, rate) - energy_F(orc, Fm, kst, ksh, rate)) / (2 * h)
    assert np.max(np.abs(P - fd)) <= 1e-6 * np.max(np.abs(fd))


@pytest.mark.parametrize("k", range(len(SAMPLES)))
def test_bw_projection_per_term_is_eigen_clamp(orc, k):
    F = SAMPLES[k]
    rate = 100.0
    f = F.T.ravel()
    # stretch only: block-diagonal in (f_u, f_v) -> clamp of the exact Hessian
    _, Hs = orc.bw_dpdf(f, 1.0, 0.0, rate)
    ref = eig_clamp(fd_hessian_F(orc, F, 1.0, 0.0, rate))
    assert np.max(np.abs(Hs - ref)) <= 1e-6 * np.max(np.abs(ref))
    # shear only
    _, Hh = orc.bw_dpdf(f, 0.0, 1.0, rate)
    ref = eig_clamp(fd_hessian_F(orc, F, 0.0, 1.0, rate))
    assert np.max(np.abs(Hh - ref)) <= 1e-6 * np.max(np.abs(ref))
    # the element Hessian is the k-weighted sum of the two projected terms
    _, Hb = orc.bw_dpdf(f, 2.0, 3.0, rate)
    assert np.max(np.abs(Hb - (2.0 * Hs + 3.0 * Hh))) <= 1e-12 * np.max(np.abs(Hb))
    assert np.linalg.eigvalsh(Hb).min() >= -1e-9 * np.abs(Hb).max()


def test_planted_shear_projection_fault_is_caught(orc):
    F = SAMPLES[0]
    ref = eig_clamp(fd_hessian_F(orc, F, 0.0, 1.0, 100.0))
    orc.set_fault(4)
    try:
        _, Hh = orc.bw_dpdf(F.T.ravel(), 0.0, 1.0, 100.0)
    finally:
        orc.set_fault(0)
    assert np.max(np.abs(Hh - ref)) > 1e-3 * np.max(np.abs(ref))


def test_bw_element_gradient_and_hessian(orc, pkg):
    s = pkg.scenes.scene_cloth_tiny(N=6, n_pairs=0)
    vals, rows, cols, grad = orc.tri_bw(s)
    assert np.all(rows <= cols)
    # gradient == dt^2 dE/dx by central differences of the summed energy
    def E(x):
        tot = 0.0
        for t in range(s.n_tris):
            tot += orc.tri_bw_energy(x[s.tris[t]].ravel(), s.dm_inv2[t].ravel(), s.area[t], s.k_stretch,
                                     s.k_shear, s.strain_rate)
        return tot
    rng = np.random.default_rng(0)
    for v in rng.choice(s.n, 6, replace=False):
        for d in range(3):
            h = 1e-7
            xp, xm = s.x.copy(), s.x.copy()
            xp[v, d] += h
            xm[v, d] -= h
            fd = (E(xp) - E(xm)) / (2 * h) * s.dt ** 2
            assert abs(grad[3 * v + d] - fd) <= 1e-5 * max(abs(fd), np.abs(grad).max() * 1e-3)
    # per triangle: symmetric PSD 9x9 with translations in the null space
    for t in range(0, s.n_tris, 7):
        H = np.zeros((9, 9))
        ids = list(s.tris[t])
        for k in range(6):
            r, c = rows[6 * t + k], cols[6 * t + k]
            a, b = ids.index(r), ids.index(c)
            blk = vals[6 * t + k].reshape(3, 3).T
            H[3 * a:3 * a + 3, 3 * b:3 * b + 3] = blk
            H[3 * b:3 * b + 3, 3 * a:3 * a + 3] = blk.T
        assert np.linalg.eigvalsh(H).min() >= -1e-9 * np.abs(H).max()
        T = np.tile(np.eye(3), (3, 1))
        assert np.max(np.abs(H @ T)) <= 1e-9 * np.abs(H).max()


def test_bw_hessian_equals_fd_where_nothing_is_clamped(orc):
    """Uniform stretch (I6 = 0, |f_u|,|f_v| > 1): no eigenvalue is clamped, so
    the projected element Hessian is the exact one."""
    class S:
        pass
    s = S()
    rest = np.array([[0, 0, 0], [1e-3, 0, 0], [0, 1e-3, 0]], float)
    s.x = rest * np.array([1.2, 1.1, 1.0])
    s.n, s.n_tris = 3, 1
    s.tris = np.array([[0, 1, 2]], np.int32)
    s.dm_inv2 = np.linalg.inv(np.array([[1e-3, 0], [0, 1e-3]]))[None]
    s.area = np.array([0.5e-6 * 1e-3])
    s.k_stretch, s.k_shear, s.strain_rate, s.dt = 3.3e5, 1e6, 100.0, 0.01
    vals, rows, cols, g0 = orc.tri_bw(s)
    H = np.zeros((9, 9))
    for k in range(6):
        a, b = rows[k], cols[k]
        blk = vals[k].reshape(3, 3).T
        H[3 * a:3 * a + 3, 3 * b:3 * b + 3] = blk
        H[3 * b:3 * b + 3, 3 * a:3 * a + 3] = blk.T
    J = np.zeros((9, 9))
    h = 1e-9
    x0 = s.x.copy()
    for q in range(9):
        s.x = x0.copy(); s.x[q // 3, q % 3] += h
        gp = orc.tri_bw(s)[3]
        s.x = x0.copy(); s.x[q // 3, q % 3] -= h
        gm = orc.tri_bw(s)[3]
        J[:, q] = (gp - gm) / (2 * h)
    s.x = x0
    assert np.max(np.abs(H - J)) <= 1e-5 * np.abs(H).max()


# ---------------------------------------------------------------- bending
def test_quad_bending_q_equilateral_closed_form(orc):
    rest = np.array([[0, 0, 0], [1, 0, 0], [0.5, np.sqrt(3) / 2, 0], [0.5, -np.sqrt(3) / 2, 0]])
    Q = orc.quad_bending_q(rest, np.array([[0, 1]]), np.array([[2, 3]])).reshape(4, 4)
    sgn = np.array([1.0, 1.0, -1.0, -1.0])
    assert np.allclose(Q, 8.0 / np.sqrt(3.0) * np.outer(sgn, sgn), rtol=1e-13, atol=1e-13)


def test_quad_bending_invariants(orc, pkg):
    s = pkg.scenes.scene_cloth_tiny(N=8, n_pairs=0)
    Q = orc.quad_bending_q(s.rest, s.edges, s.adj).reshape(-1, 4, 4)
    interior = s.adj[:, 1] >= 0
    assert np.all(Q[~interior] == 0)
    Qi = Q[interior]
    assert np.max(np.abs(Qi @ np.ones(4))) <= 1e-10 * np.abs(Qi).max()        # Q 1 = 0
    assert np.max(np.abs(Qi - Qi.transpose(0, 2, 1))) == 0
    w = np.linalg.eigvalsh(Qi)
    assert w.min() >= -1e-10 * np.abs(w).max()
    assert np.all(np.sum(w > 1e-10 * np.abs(w).max(), axis=1) == 1)            # rank 1
    # flat rest state: the cotangent operator annihilates planar positions
    ids = np.concatenate([s.edges, s.adj], axis=1)[interior]
    Kx = np.einsum("hab,hbd->had", Qi, s.rest[ids])
    assert np.max(np.abs(Kx)) <= 1e-9 * np.abs(Qi).max() * 1e-3


def test_quad_bending_gradient_hessian(orc, pkg):
    s = pkg.scenes.scene_cloth_tiny(N=8, n_pairs=0)
    Q = orc.quad_bending_q(s.rest, s.edges, s.adj)
    vals, rows, cols, grad = orc.quad_bending(s, Q)
    coefs = np.linalg.norm(s.rest[s.edges[:, 0]] - s.rest[s.edges[:, 1]], axis=1) * s.k_bend * s.dt ** 2
    def E(x):
        ids = np.concatenate([s.edges, s.adj], axis=1)
        tot = 0.0
        for h in range(s.n_hinges):
            if s.adj[h, 1] < 0:
                continue
            xs = x[ids[h]]
            tot += 0.5 * coefs[h] * np.sum(Q[h].reshape(4, 4) * (xs @ xs.T))
        return tot
    rng = np.random.default_rng(1)
    for v in rng.choice(s.n, 5, replace=False):
        for d in range(3):
            hh = 1e-6
            xp, xm = s.x.copy(), s.x.copy()
            xp[v, d] += hh
            xm[v, d] -= hh
            fd = (E(xp) - E(xm)) / (2 * hh)
            assert abs(grad[3 * v + d] - fd) <= 1e-6 * np.abs(grad).max()
    # rotation invariance of the energy
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    assert abs(E(s.x @ Rz.T) - E(s.x)) <= 1e-10 * abs(E(s.x))
    # boundary hinges write ten zero blocks at (0, 0); interior blocks are coef * Q_ab * I
    for h in range(s.n_hinges):
        blk = vals[10 * h:10 * h + 10].reshape(10, 3, 3)
        if s.adj[h, 1] < 0:
            assert np.all(blk == 0) and np.all(rows[10 * h:10 * h + 10] == 0) and np.all(cols[10 * h:10 * h + 10] == 0)
        else:
            assert np.all(rows[10 * h:10 * h + 10] <= cols[10 * h:10 * h + 10])
            off = blk - np.eye(3)[None] * blk[:, :1, :1]
            assert np.all(off == 0)
