"""Pins of the oracle's MAS coarse mode 2 (oracle/mpnn.py: mas_vcycle_apply),
the symmetric two-grid cycle on level 0 around the level-1 cycle.

Built as dense operators on a small scene: the cycle's error propagation
must factor as (I - S A)(I - C A)(I - S A) with S = omega M0 (the level-0
domain solves) and C = P V1 P^T (the level-1 cycle, prolonged) -- the
textbook form of a pre-smooth / coarse-correct / post-smooth cycle -- and the
preconditioner must be symmetric positive definite for the default omega.
A wrong order of the steps, a residual taken against the wrong vector or a
misplaced omega breaks the factorisation."""
import numpy as np
import pytest

from oracle import mpnn


@pytest.fixture(scope="module")
def setup(pkg, orc):
    s = pkg.scenes.tet_grid_scene("vc_grid9", 9, E_loguniform=(1e4, 1e9), n_pairs=60)
    vals, rows, cols = orc.assemble(s)
    B = 16
    m = orc.Mas(s, B)
    m.build(vals, rows, cols, 0, s.pairs)
    parts = mpnn.MasParts(m)
    assert parts.L > 2 and parts.cnt[1] > B
    A = mpnn.bcoo_to_csr(vals, rows, cols, s.n)
    graph = mpnn.coarse_graph(A, parts.map1, parts.cnt[1])
    A1f = mpnn.float_operator(graph)
    A2 = mpnn.level2_operator(parts, graph)
    yield s, parts, A, graph, A1f, A2
    m.close()


def _columns(f, n3):
    M = np.zeros((n3, n3))
    for j in range(n3):
        e = np.zeros(n3)
        e[j] = 1.0
        M[:, j] = f(e)
    return M


def test_two_grid_error_propagation_factorises(setup):
    s, parts, A, graph, A1f, A2 = setup
    n3 = 3 * s.n
    om = 0.5
    AP = (A @ graph.P).tocsr()  # exact A P: the factorisation holds exactly
    Bv = _columns(lambda r: mpnn.mas_vcycle_apply(parts, A1f, A, AP, r, A2=A2, omega=om), n3)
    M0 = _columns(lambda r: mpnn.level0_apply(parts, r), n3)
    P = graph.P.toarray()
    c1 = parts.cnt[1]
    V1 = np.zeros((3 * c1, 3 * c1))
    for j in range(3 * c1):
        e = np.zeros(3 * c1)
        e[j] = 1.0
        V1[:, j] = mpnn.level1_cycle(parts, A1f, e.reshape(-1, 3), A2).reshape(-1)
    Ad = A.toarray()
    I = np.eye(n3)
    S = om * M0
    C = P @ V1 @ P.T
    E = (I - S @ Ad) @ (I - C @ Ad) @ (I - S @ Ad)
    err = np.abs((I - Bv @ Ad) - E).max()
    assert err <= 1e-8 * np.abs(E).max(), err


def test_vcycle_is_spd_and_symmetric_with_exact_ap(setup):
    s, parts, A, graph, A1f, A2 = setup
    n3 = 3 * s.n
    AP = (A @ graph.P).tocsr()
    Bv = _columns(lambda r: mpnn.mas_vcycle_apply(parts, A1f, A, AP, r, A2=A2, omega=0.5), n3)
    asym = np.abs(Bv - Bv.T).max() / np.abs(Bv).max()
    assert asym <= 1e-10, asym
    assert np.linalg.eigvalsh(0.5 * (Bv + Bv.T)).min() > 0


def test_vcycle_beats_the_additive_level0_in_pcg(setup):
    """Fewer PCG iterations than the mode-1 apply (additive level 0) on the
    same hierarchy, with the float32 A P the library stores."""
    s, parts, A, graph, A1f, A2 = setup
    APf = mpnn.ap_operator(A, parts.map1, parts.cnt[1])
    assert APf.nnz < A.nnz  # A P is sparser than A

    def pcg(prec, tol=1e-8):
        b = s.rhs
        x = np.zeros_like(b)
        r = b.copy()
        z = prec(r)
        p = z.copy()
        rz = r @ z
        for k in range(1, 2000):
            Ap = A @ p
            a = rz / (p @ Ap)
            x += a * p
            r -= a * Ap
            if np.linalg.norm(r) <= tol * np.linalg.norm(b):
                return k
            z = prec(r)
            rz_new = r @ z
            p = z + (rz_new / rz) * p
            rz = rz_new
        return 2000

    k1 = pcg(lambda r: mpnn.mas_cycle_apply(parts, A1f, r, A2=A2))
    k2 = pcg(lambda r: mpnn.mas_vcycle_apply(parts, A1f, A, APf, r, A2=A2, omega=0.5))
    print(f"level-1 cycle {k1} iterations, two-grid on level 0 {k2}")
    assert k2 < k1


def test_half_level0_storage_keeps_spd_and_iterations(setup):
    """The FP16 storage of the level-0 inverses (unit-diagonal scaling, then
    float16) keeps every level-0 domain matrix positive definite, perturbs
    each entry by at most 2^-11 of d_i d_j, and leaves the two-grid PCG
    iteration count within one of the float64 inverses."""
    s, parts, A, graph, A1f, A2 = setup
    B = parts.B
    n0 = parts.off[1] // B
    inv = parts.inv
    h = mpnn.half_level0_inverse(inv, n0)
    assert np.array_equal(h[n0:], inv[n0:])
    for k in range(n0):
        real = np.diag(inv[k]) > 0
        Mk, Hk = inv[k][np.ix_(real, real)], h[k][np.ix_(real, real)]
        d = np.sqrt(np.diag(Mk))
        # float16 keeps 11 significant bits of each |C_ij| <= 1 (plus the
        # float32 rounding of the scaling)
        assert np.all(np.abs(Hk - Mk) <= (2.0 ** -11 + 1e-6) * np.outer(d, d))
        assert np.linalg.eigvalsh(Hk).min() > 0
    APf = mpnn.ap_operator(A, parts.map1, parts.cnt[1])

    def pcg(prec, tol=1e-8):
        b = s.rhs
        x = np.zeros_like(b)
        r = b.copy()
        z = prec(r)
        p = z.copy()
        rz = r @ z
        for k in range(1, 2000):
            Ap = A @ p
            a = rz / (p @ Ap)
            x += a * p
            r -= a * Ap
            if np.linalg.norm(r) <= tol * np.linalg.norm(b):
                return k
            z = prec(r)
            rz_new = r @ z
            p = z + (rz_new / rz) * p
            rz = rz_new
        return 2000

    k_full = pcg(lambda r: mpnn.mas_vcycle_apply(parts, A1f, A, APf, r, A2=A2, omega=0.4))
    parts.inv = h
    try:
        k_half = pcg(lambda r: mpnn.mas_vcycle_apply(parts, A1f, A, APf, r, A2=A2, omega=0.4))
    finally:
        parts.inv = inv
    print(f"two-grid PCG: float64 level-0 inverses {k_full} iterations, fp16 storage {k_half}")
    assert k_half <= k_full + 1
