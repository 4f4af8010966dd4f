"""Pins the oracle's element Hessians against closed forms and finite
differences (independent of the oracle's own formulas).

* Stable Neo-Hookean: Psi(F) = mu/2 (|F|^2-3) - mu (J-1) + lam/2 (J-1)^2
  (femEnergy.cu:432-459 is its PK1).  The gradient must equal the FD of the
  energy; away from the projection the Hessian must equal the FD of the
  gradient; the projected 9x9 must equal numpy's eigh-clamp of the analytic one;
  at rest the twist modes are exactly zero; rigid translations are in the
  null space.
* GIPC rank-2 PT barrier: E(I5) = kappa dHat^2 (I5-1)^2 ln(I5)^2, I5 = s^2,
  s = dist/sqrt(dHat) (GIPC.cu:4733-4984).  lambda0 must be d2E/ds2 (FD), the
  gradient the FD of E(x) with dist from numpy's point-plane formula, the
  Hessian rank one lambda0 g g^T with g the FD gradient of s.
"""
import numpy as np
import pytest


def tet_scene_one(pkg, F=None, seed=0, mu=1.3e4, lam=8.0e4):
    """A single tet with prescribed deformation gradient F."""
    rng = np.random.default_rng(seed)
    rest = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.01
    rest += rng.uniform(-1e-3, 1e-3, size=rest.shape)
    Dm = np.stack([rest[1] - rest[0], rest[2] - rest[0], rest[3] - rest[0]], axis=1)
    if F is None:
        F = np.eye(3) + 0.15 * rng.standard_normal((3, 3))
    x = (F @ (rest - rest[0]).T).T + 0.3
    return x, np.linalg.inv(Dm), abs(np.linalg.det(Dm)) / 6.0, mu, lam


def assemble12(vals, rows, cols, gid):
    """Dense 12x12 from the 10 upper blocks of one element."""
    H = np.zeros((12, 12))
    loc = {g: a for a, g in enumerate(gid)}
    for v, r, c in zip(vals, rows, cols):
        M = v.reshape(3, 3).T  # column-major
        a, b = loc[r], loc[c]
        H[3 * a:3 * a + 3, 3 * b:3 * b + 3] += M
        if a != b:
            H[3 * b:3 * b + 3, 3 * a:3 * a + 3] += M.T
    return H


def oracle_tet(orc, x, dminv, vol, mu, lam, dt=0.01, gid=(0, 1, 2, 3)):
    import ctypes
    L = orc.lib()
    n = 4
    xs = np.zeros((max(gid) + 1, 3))
    for a, g in enumerate(gid):
        xs[g] = x[a]
    tets = np.array([gid], np.int32)
    vals = np.zeros((10, 9))
    rows = np.zeros(10, np.int32)
    cols = np.zeros(10, np.int32)
    grad = np.zeros(xs.size)
    P = ctypes.POINTER
    L.orc_tet_snh(1, xs.ctypes.data_as(P(ctypes.c_double)), tets.ctypes.data_as(P(ctypes.c_int)),
                  np.ascontiguousarray(dminv).ctypes.data_as(P(ctypes.c_double)),
                  np.array([vol]).ctypes.data_as(P(ctypes.c_double)), np.array([mu]).ctypes.data_as(P(ctypes.c_double)),
                  np.array([lam]).ctypes.data_as(P(ctypes.c_double)), dt, 0, grad.ctypes.data_as(P(ctypes.c_double)),
                  vals.ctypes.data_as(P(ctypes.c_double)), rows.ctypes.data_as(P(ctypes.c_int)),
                  cols.ctypes.data_as(P(ctypes.c_int)))
    g12 = np.concatenate([grad[3 * g:3 * g + 3] for g in gid])
    return assemble12(vals, rows, cols, list(gid)), g12, (vals, rows, cols)


def energy(orc, x, dminv, mu, lam):
    import ctypes
    P = ctypes.POINTER(ctypes.c_double)
    x12 = np.ascontiguousarray(x.reshape(-1))
    return orc.lib().orc_tet_snh_energy(x12.ctypes.data_as(P), np.ascontiguousarray(dminv).ctypes.data_as(P), mu, lam)


def test_snh_gradient_is_fd_of_energy(pkg, orc):
    x, dminv, vol, mu, lam = tet_scene_one(pkg)
    dt = 0.01
    _, g, _ = oracle_tet(orc, x, dminv, vol, mu, lam, dt)
    h = 1e-7
    fd = np.zeros(12)
    for k in range(12):
        xp, xm = x.copy(), x.copy()
        xp.flat[k] += h
        xm.flat[k] -= h
        fd[k] = dt * dt * vol * (energy(orc, xp, dminv, mu, lam) - energy(orc, xm, dminv, mu, lam)) / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-6 * np.max(np.abs(fd))


def test_snh_hessian_is_fd_of_gradient_when_unprojected(pkg, orc):
    # mild stretch: every eigenvalue of dP/dF is positive, projection inactive
    F = np.diag([1.05, 1.02, 0.99])
    x, dminv, vol, mu, lam = tet_scene_one(pkg, F=F, seed=3)
    dt = 0.01
    H, _, _ = oracle_tet(orc, x, dminv, vol, mu, lam, dt)
    h = 1e-8
    fd = np.zeros((12, 12))
    for k in range(12):
        xp, xm = x.copy(), x.copy()
        xp.flat[k] += h
        xm.flat[k] -= h
        fd[:, k] = (oracle_tet(orc, xp, dminv, vol, mu, lam, dt)[1] - oracle_tet(orc, xm, dminv, vol, mu, lam, dt)[1]) / (2 * h)
    assert np.max(np.abs(H - fd)) <= 1e-5 * np.max(np.abs(fd))


@pytest.mark.parametrize("seed", range(6))
def test_snh_projection_equals_eigh_clamp(orc, seed):
    import ctypes
    rng = np.random.default_rng(seed)
    F = np.eye(3) + 0.4 * rng.standard_normal((3, 3))  # often compressed/inverted: projection active
    F9 = np.ascontiguousarray(F.T.reshape(-1))  # column-major vec
    H = np.zeros(81)
    Hp = np.zeros(81)
    P = ctypes.POINTER(ctypes.c_double)
    orc.lib().orc_snh_dpdf(F9.ctypes.data_as(P), 1.0e4, 9.0e4, H.ctypes.data_as(P), Hp.ctypes.data_as(P))
    H = H.reshape(9, 9)
    w, V = np.linalg.eigh(0.5 * (H + H.T))
    ref = (V * np.maximum(w, 0)) @ V.T
    assert np.max(np.abs(Hp.reshape(9, 9) - ref)) <= 1e-9 * np.max(np.abs(ref))
    # the analytic Hessian itself is the FD of PK1 = mu F + (lam(J-1)-mu) cof(F)
    def pk1(Fm):
        J = np.linalg.det(Fm)
        cof = J * np.linalg.inv(Fm).T
        return 1.0e4 * Fm + (9.0e4 * (J - 1) - 1.0e4) * cof
    h = 1e-7
    fd = np.zeros((9, 9))
    for q in range(9):
        i, j = q % 3, q // 3
        Fp, Fm_ = F.copy(), F.copy()
        Fp[i, j] += h
        Fm_[i, j] -= h
        fd[:, q] = ((pk1(Fp) - pk1(Fm_)) / (2 * h)).T.reshape(-1)
    assert np.max(np.abs(H - fd)) <= 1e-5 * np.max(np.abs(fd))


def test_snh_rest_state_modes(orc):
    import ctypes
    P = ctypes.POINTER(ctypes.c_double)
    mu, lam = 2.0, 7.0
    H = np.zeros(81)
    F9 = np.eye(3).reshape(-1).copy()
    orc.lib().orc_snh_dpdf(F9.ctypes.data_as(P), mu, lam, H.ctypes.data_as(P), None)
    w = np.sort(np.linalg.eigvalsh(H.reshape(9, 9)))
    # at F = I: twist modes mu - mu = 0 (rotations), flip modes 2 mu, scaling
    # block 2 mu I + (lam - mu) 1 1^T -> 2 mu (twice) and 3 lam - mu
    np.testing.assert_allclose(w, [0, 0, 0, 2 * mu, 2 * mu, 2 * mu, 2 * mu, 2 * mu, 3 * lam - mu], atol=1e-12)


def test_snh_translation_in_null_space(pkg, orc):
    x, dminv, vol, mu, lam = tet_scene_one(pkg, seed=5)
    H, _, _ = oracle_tet(orc, x, dminv, vol, mu, lam)
    for t in np.eye(3):
        assert np.max(np.abs(H @ np.tile(t, 4))) <= 1e-10 * np.max(np.abs(H))
    assert np.min(np.linalg.eigvalsh(H)) >= -1e-10 * np.max(np.abs(H))


def test_snh_upper_block_writer(pkg, orc):
    # vertex ids out of order: blocks must be stored at (min, max) transposed
    x, dminv, vol, mu, lam = tet_scene_one(pkg, seed=7)
    gid = (9, 2, 5, 0)
    H, _, (vals, rows, cols) = oracle_tet(orc, x, dminv, vol, mu, lam, gid=gid)
    assert np.all(rows <= cols)
    H0, _, _ = oracle_tet(orc, x, dminv, vol, mu, lam, gid=(0, 1, 2, 3))
    np.testing.assert_allclose(H, H0, rtol=0, atol=1e-12 * np.max(np.abs(H0)))


# ----------------------------------------------------------------- barrier
def test_barrier_lambda0_is_second_derivative(orc):
    L = orc.lib()
    dh, kappa = 1e-6, 1e6
    for I5 in (1e-4, 0.01, 0.1, 0.3, 0.5, 0.8, 0.95):
        s = np.sqrt(I5)
        h = 1e-5 * s
        E = lambda ss: L.orc_barrier_energy(ss * ss, dh, kappa)  # noqa: E731
        fd = (E(s + h) - 2 * E(s) + E(s - h)) / (h * h)
        lam0 = L.orc_barrier_lambda0(I5, dh, kappa)
        assert abs(lam0 - fd) <= 1e-4 * abs(fd), (I5, lam0, fd)
        assert lam0 > 0


def _pt_geometry(rng, d):
    t = rng.standard_normal((3, 3)) * 0.01
    n = np.cross(t[1] - t[0], t[2] - t[0])
    n /= np.linalg.norm(n)
    bary = rng.uniform(0.2, 1.0, 3)
    bary /= bary.sum()
    p = bary @ t + d * n
    return np.vstack([p, t])


def _dist(P):
    b = np.cross(P[2] - P[1], P[3] - P[1])
    return abs(np.dot(P[0] - P[1], b)) / np.linalg.norm(b)


@pytest.mark.parametrize("seed", range(4))
def test_barrier_gradient_and_rank_one_hessian(orc, seed):
    import ctypes
    rng = np.random.default_rng(seed)
    dhat = 1e-3
    dh, kappa = dhat * dhat, 1e6
    P = _pt_geometry(rng, dhat * rng.uniform(0.2, 0.9))
    pairs = np.array([[-1, 1, 2, 3]], np.int32)  # point 0, triangle (1,2,3)
    vals = np.zeros((16, 9))
    rows = np.zeros(16, np.int32)
    cols = np.zeros(16, np.int32)
    grad = np.zeros(12)
    x = np.ascontiguousarray(P.reshape(-1))
    Pd = ctypes.POINTER(ctypes.c_double)
    Pi = ctypes.POINTER(ctypes.c_int)
    orc.lib().orc_barrier_pt(1, x.ctypes.data_as(Pd), pairs.ctypes.data_as(Pi), None, dh, kappa, 0,
                             grad.ctypes.data_as(Pd), vals.ctypes.data_as(Pd), rows.ctypes.data_as(Pi),
                             cols.ctypes.data_as(Pi))

    def E(Pm):
        I5 = _dist(Pm) ** 2 / dh
        return orc.lib().orc_barrier_energy(I5, dh, kappa)

    def s_of(Pm):
        return _dist(Pm) / np.sqrt(dh)

    h = 1e-9
    fdg = np.zeros(12)
    gs = np.zeros(12)
    for k in range(12):
        Pp, Pm = P.copy(), P.copy()
        Pp.flat[k] += h
        Pm.flat[k] -= h
        fdg[k] = (E(Pp) - E(Pm)) / (2 * h)
        gs[k] = (s_of(Pp) - s_of(Pm)) / (2 * h)
    assert np.max(np.abs(grad - fdg)) <= 1e-5 * np.max(np.abs(fdg))
    H = np.zeros((12, 12))
    for v, r, c in zip(vals, rows, cols):
        M = v.reshape(3, 3).T
        H[3 * r:3 * r + 3, 3 * c:3 * c + 3] += M
        if r != c:
            H[3 * c:3 * c + 3, 3 * r:3 * r + 3] += M.T
    lam0 = orc.lib().orc_barrier_lambda0(_dist(P) ** 2 / dh, dh, kappa)
    ref = lam0 * np.outer(gs, gs)
    assert np.max(np.abs(H - ref)) <= 1e-5 * np.max(np.abs(ref))
    assert np.linalg.matrix_rank(H, tol=1e-9 * np.max(np.abs(H))) == 1
    # lower blocks of the 4x4 pattern are moved to (col,row) and zeroed
    low = [k for k in range(16) if k // 4 > k % 4]
    assert np.all(vals[low] == 0) and np.all(rows <= cols)


def test_planted_barrier_fault_is_caught(orc):
    """A wrong sign of the log term in lambda0 breaks the d2E/ds2 check."""
    L = orc.lib()
    orc.set_fault(3)
    try:
        I5, dh, kappa = 0.3, 1e-6, 1e6
        s = np.sqrt(I5)
        h = 1e-5 * s
        E = lambda ss: L.orc_barrier_energy(ss * ss, dh, kappa)  # noqa: E731
        fd = (E(s + h) - 2 * E(s) + E(s - h)) / (h * h)
        assert abs(L.orc_barrier_lambda0(I5, dh, kappa) - fd) > 1e-2 * abs(fd)
    finally:
        orc.set_fault(0)
