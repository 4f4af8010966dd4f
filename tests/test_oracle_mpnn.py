"""Pins of the learned-coarse-level oracle (oracle/mpnn.py) on the CPU."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import mpnn


@pytest.fixture(scope="module")
def tiny_setup(orc, scene_tiny):
    s = scene_tiny
    vals, rows, cols = orc.assemble(s)
    m = orc.Mas(s, 16)
    m.build(vals, rows, cols, 0, s.pairs)
    parts = mpnn.MasParts(m)
    A = mpnn.bcoo_to_csr(vals, rows, cols, s.n)
    graph = mpnn.coarse_graph(A, parts.map1, parts.cnt[1])
    yield s, A, m, parts, graph
    m.close()


def _random_spd_graph(nodes, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(3 * nodes, 3 * nodes, density=0.08, random_state=seed)
    A = (M @ M.T + sp.eye(3 * nodes) * 3.0).tocsr()
    return A, rng


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-30, 30, 4000),
                        [1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 0.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mpnn.bf16(x), ref)


@pytest.mark.parametrize("linear", [False, True])
def test_gradient_matches_finite_differences(linear):
    A, rng = _random_spd_graph(12, 1)
    C, L = 8, 2
    p = mpnn.init_params(C, L, seed=3)
    p.w_out *= 10.0  # a correction large enough for every path to matter
    samples = [(rng.standard_normal((A.shape[0], 3)), rng.standard_normal(A.shape[0])) for _ in range(3)]
    f0, g = mpnn.loss_grad(p, A, samples, linear=linear)
    v0 = p.flat()
    gflat = g.flat()
    for k in range(6):
        d = rng.standard_normal(v0.size)
        h = 1e-6
        fp, _ = mpnn.loss_grad(mpnn.Params.from_flat(v0 + h * d, C, L), A, samples, linear=linear)
        fm, _ = mpnn.loss_grad(mpnn.Params.from_flat(v0 - h * d, C, L), A, samples, linear=linear)
        fd = (fp - fm) / (2 * h)
        assert abs(fd - gflat @ d) <= 1e-6 * max(1.0, abs(fd)), (k, fd, gflat @ d)


def test_loss_minimiser_is_the_coarse_solve():
    A, rng = _random_spd_graph(10, 2)
    q = rng.standard_normal(A.shape[0])
    lmin, c = mpnn.optimal_loss(A, q)
    ref = -0.5 * q @ np.linalg.solve(A.toarray(), q) / q.size
    assert abs(lmin - ref) <= 1e-12 * abs(ref)
    for _ in range(3):
        d = rng.standard_normal(q.size) * 0.1
        # the loss is a quadratic with Hessian A / rows around its minimiser
        assert abs(mpnn.loss(c + d, A, q) - lmin - 0.5 * d @ (A @ d) / q.size) <= 1e-12


def test_network_is_positively_homogeneous():
    A, rng = _random_spd_graph(9, 4)
    p = mpnn.init_params(16, 3, seed=5)
    X = rng.standard_normal((A.shape[0], 3))
    c1, _ = mpnn.forward(p, A, X)
    c2, _ = mpnn.forward(p, A, 7.5 * X)
    np.testing.assert_allclose(c2, 7.5 * c1, rtol=1e-12, atol=1e-14)
    c3, _ = mpnn.forward(p, A, -X)
    assert np.linalg.norm(c3 + c1) > 1e-3 * np.linalg.norm(c1)  # relu: not odd


def test_level1_operator_matches_mas_galerkin(tiny_setup):
    """The diagonal domain blocks of A_1 = P^T A P are the MAS level-1
    Galerkin blocks computed by the C oracle."""
    s, A, m, parts, graph = tiny_setup
    B = m.B
    gal = mpnn._domain_dense(m.matrices(inverse=False), B)
    off = parts.off
    c1 = parts.cnt[1]
    A1 = graph.A1.toarray()
    for d in range(off[1] // B, off[2] // B):
        lo = d * B - off[1]
        hi = min(lo + B, c1)
        ref = A1[3 * lo:3 * hi, 3 * lo:3 * hi]
        got = gal[d][:3 * (hi - lo), :3 * (hi - lo)]
        assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref))
    # unit average diagonal after scaling
    dg = graph.Ahat.diagonal().reshape(-1, 3).sum(axis=1) / 3
    np.testing.assert_allclose(dg, 1.0, rtol=1e-12)


def test_mas_parts_match_the_c_oracle(tiny_setup):
    s, A, m, parts, graph = tiny_setup
    r = np.random.default_rng(6).standard_normal(3 * s.n)
    z, zc = parts.apply(r)
    ref = m.apply(r)
    assert np.max(np.abs(z - ref)) <= 1e-11 * np.max(np.abs(ref))
    assert zc.size == 3 * parts.cnt[1]


def test_loss_is_the_a_norm_error_of_the_correction(tiny_setup):
    """2 rows loss = ||e - M r - P c||_A^2 - ||e - M r||_A^2 for the
    solution e = A^-1 r of the residual: the loss trains on
    residual/solution pairs without forming e."""
    s, A, m, parts, graph = tiny_setup
    rng = np.random.default_rng(7)
    e = rng.standard_normal(3 * s.n)
    r = A @ e  # an exact residual/solution pair
    X, qhat, z = mpnn.sample(A, graph, parts, r)
    sc = np.sqrt(np.mean((np.repeat(graph.s, 3) * (graph.P.T @ r)) ** 2))
    chat = rng.standard_normal(qhat.size) / sc * 1e-3
    c = np.repeat(graph.s, 3) * chat * sc
    err0 = e - z
    err1 = err0 - graph.P @ c
    lhs = err1 @ (A @ err1) - err0 @ (A @ err0)
    rhs = 2 * qhat.size * mpnn.loss(chat, graph.Ahat, qhat) * sc * sc
    assert abs(lhs - rhs) <= 1e-8 * abs(lhs)


def test_adam_reduces_the_loss(tiny_setup):
    s, A, m, parts, graph = tiny_setup
    rng = np.random.default_rng(8)
    samples = []
    for _ in range(12):
        X, qhat, _ = mpnn.sample(A, graph, parts, rng.standard_normal(3 * s.n))
        samples.append((X, qhat))
    p = mpnn.init_params(16, 2, seed=1)
    p2, hist = mpnn.adam(p, graph.Ahat, samples, epochs=30, batch=4, lr=3e-3)
    l0, _ = mpnn.loss_grad(p, graph.Ahat, samples)
    l1, _ = mpnn.loss_grad(p2, graph.Ahat, samples)
    lopt = np.mean([mpnn.optimal_loss(graph.Ahat, q)[0] for _, q in samples])
    assert l1 < l0 and l1 < 0.2 * lopt, (l0, l1, lopt)  # captures a share of the optimum
