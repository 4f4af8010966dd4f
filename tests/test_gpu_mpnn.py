"""Learned coarse level on the GPU against the numpy oracle (oracle/mpnn.py):
graph build, BF16-MFMA forward pass, gradients and Adam training."""
import numpy as np
import pytest

from oracle import mpnn

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def level1(orc, scene_tiny):
    """Level-1 operator of the tiny stiff scene (B = 16) from the oracle."""
    s = scene_tiny
    vals, rows, cols = orc.assemble(s)
    m = orc.Mas(s, 16)
    m.build(vals, rows, cols, 0, s.pairs)
    parts = mpnn.MasParts(m)
    A = mpnn.bcoo_to_csr(vals, rows, cols, s.n)
    graph = mpnn.coarse_graph(A, parts.map1, parts.cnt[1])
    rng = np.random.default_rng(11)
    samples = []
    for _ in range(12):
        X, q, _ = mpnn.sample(A, graph, parts, rng.standard_normal(3 * s.n))
        samples.append((X, X[:, 2].copy()))
    yield graph, samples
    m.close()


def _data(pkg, ctx, graph, samples):
    v, r, c = mpnn.coarse_upper_bcoo(graph.A1, graph.n)
    d = pkg.learned.MpnnData.from_bcoo(ctx, graph.n, v, r, c)
    d.add(np.stack([X for X, _ in samples]), np.stack([q for _, q in samples]))
    return d


def _ahat32(d):
    """Ahat as built on the device (float32 blocks), as a scipy matrix."""
    import scipy.sparse as sp
    rp, col, blk, sc = d.graph()
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    R = (3 * rows[:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2)
    Cc = (3 * col[:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1)
    return sp.csr_matrix((blk.astype(np.float64).reshape(-1), (R.reshape(-1), Cc.reshape(-1))), shape=(3 * n, 3 * n)), sc


def test_graph_matches_oracle(pkg, gpu_ctx, level1):
    graph, samples = level1
    d = _data(pkg, gpu_ctx, graph, samples)
    A32, sc = _ahat32(d)
    np.testing.assert_allclose(sc, graph.s, rtol=1e-6)
    D = (A32 - graph.Ahat).toarray()
    assert np.abs(D).max() <= 1e-6 * np.abs(graph.Ahat.toarray()).max()
    # columns sorted inside every row, both directions present
    rp, col, _, _ = d.graph()
    for i in range(rp.size - 1):
        assert np.all(np.diff(col[rp[i]:rp[i + 1]]) > 0)
    assert col.size == graph.Ahat.nnz // 9
    d.close()


@pytest.mark.parametrize("linear", [False, True])
def test_forward_matches_oracle(pkg, gpu_ctx, level1, linear):
    graph, samples = level1
    d = _data(pkg, gpu_ctx, graph, samples)
    A32, _ = _ahat32(d)
    net = pkg.learned.Mpnn(gpu_ctx, layers=3, linear=linear, seed=5)
    p = mpnn.Params.from_flat(net.get_params().astype(np.float64), 64, 3)
    p.w_out *= 30.0  # make the output O(1)
    net.set_params(p.flat())
    got = net.forward(d, 0, 4).reshape(4, -1)
    for k in range(4):
        ref, _ = mpnn.forward(p, A32, samples[k][0].astype(np.float32).astype(np.float64), bf16_ops=True,
                              linear=linear)
        err = np.linalg.norm(got[k] - ref) / np.linalg.norm(ref)
        assert err <= 5e-3, (k, err)
    net.close()
    d.close()


@pytest.mark.parametrize("linear", [False, True])
def test_gradient_matches_oracle(pkg, gpu_ctx, level1, linear):
    """GPU forward/backward kernels against the oracle's gradient with the
    same bf16 rounding of the GEMM operands (the oracle's gradient is itself
    pinned by finite differences).  Where the two round a value to different
    bf16 neighbours the difference is one bf16 ulp (2^-8), and backpropagation
    amplifies it layer by layer: two layers at realistic output scale keep it
    near 1e-2."""
    graph, samples = level1
    d = _data(pkg, gpu_ctx, graph, samples)
    A32, _ = _ahat32(d)
    L = 2
    net = pkg.learned.Mpnn(gpu_ctx, layers=L, linear=linear, seed=7)
    p = mpnn.Params.from_flat(net.get_params().astype(np.float64), 64, L)
    loss, g = net.loss_grad(d, 0, 6)
    sm = [(X.astype(np.float32).astype(np.float64), q.astype(np.float32).astype(np.float64)) for X, q in samples[:6]]
    lref, gref = mpnn.loss_grad(p, A32, sm, bf16_ops=True, linear=linear)
    assert abs(loss - lref) <= 2e-3 * abs(lref), (loss, lref)
    gr = gref.flat()
    cos = g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr))
    assert cos >= 0.999, cos
    C = 64
    for lo, hi in [(0, 3 * C), (3 * C, 3 * C + L * 2 * C * C), (3 * C + L * 2 * C * C, g.size)]:
        e = np.linalg.norm(g[lo:hi] - gr[lo:hi]) / np.linalg.norm(gr[lo:hi])
        assert e <= 5e-2, (lo, e)
    net.close()
    d.close()


@pytest.mark.parametrize("linear", [False, True])
def test_three_layer_gradient_is_a_descent_direction(pkg, gpu_ctx, level1, linear):
    """Three layers: the device gradient points where the oracle's does
    (cosine; the bf16 ulp differences grow with depth, so per-block errors
    are not compared), and a step along -gradient lowers the device's own
    loss.  A finite-difference slope of the device loss is not used: a
    parameter step re-rounds every bf16 operand, which puts about 1e-3
    relative noise into each loss evaluation."""
    graph, samples = level1
    d = _data(pkg, gpu_ctx, graph, samples)
    A32, _ = _ahat32(d)
    net = pkg.learned.Mpnn(gpu_ctx, layers=3, linear=linear, seed=9)
    w = net.get_params().astype(np.float64)
    p = mpnn.Params.from_flat(w, 64, 3)
    l0, g = net.loss_grad(d, 0, 6)
    g = g.astype(np.float64)
    sm = [(X.astype(np.float32).astype(np.float64), q.astype(np.float32).astype(np.float64)) for X, q in samples[:6]]
    _, gref = mpnn.loss_grad(p, A32, sm, bf16_ops=True, linear=linear)
    gr = gref.flat()
    cos = g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr))
    assert cos >= 0.98, cos
    step = 0.05 * abs(l0) / np.linalg.norm(g) ** 2  # a 5% predicted decrease
    net.set_params((w - step * g).astype(np.float32))
    l1, _ = net.loss_grad(d, 0, 6)
    assert l1 < l0 - 0.02 * abs(l0), (l0, l1)
    net.close()
    d.close()


def test_training_reduces_the_loss(pkg, gpu_ctx, level1):
    graph, samples = level1
    d = _data(pkg, gpu_ctx, graph, samples)
    net = pkg.learned.Mpnn(gpu_ctx, layers=2, linear=False, seed=3)
    l0, _ = net.loss_grad(d, 0, d.size)
    hist = net.train(d, epochs=40, batch=4, lr=3e-3, seed=1)
    l1, _ = net.loss_grad(d, 0, d.size)
    lopt = np.mean([mpnn.optimal_loss(graph.Ahat, q)[0] for _, q in samples])
    print(f"loss {l0:.4e} -> {l1:.4e}; optimum {lopt:.4e}; epochs {hist[0]:.3e} .. {hist[-1]:.3e}")
    assert np.all(np.isfinite(hist))
    assert l1 < l0 and l1 <= 0.2 * lopt, (l0, l1, lopt)
    net.close()
    d.close()
