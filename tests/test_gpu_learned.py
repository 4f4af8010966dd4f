"""The learned coarse level inside the solve: training samples collected from
the plain solver, training on the GPU, flexible PCG with the correction on a
held-out scene, and the automatic fallback to plain MAS."""
import numpy as np
import pytest

from oracle import mpnn

pytestmark = pytest.mark.gpu


def _scene(pkg, seed, N=12):
    return pkg.scenes.tet_grid_scene(f"learn_grid{N}_{seed}", N, E_loguniform=(1e4, 1e9), n_pairs=N ** 3 // 5,
                                     seed_E=seed, seed_pairs=seed + 100, seed_jitter=seed, seed_rhs=seed + 7)


def _numpy_pcg_residuals(A, b, prec, k):
    """Residuals after each of the first k PCG iterations (textbook PCG)."""
    r = b.copy()
    z = prec(r)
    p = z.copy()
    rz = r @ z
    out = []
    for _ in range(k):
        Ap = A @ p
        alpha = rz / (p @ Ap)
        r = r - alpha * Ap
        out.append(r.copy())
        z = prec(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return out


def test_collected_samples_match_oracle(pkg, orc):
    s = _scene(pkg, 1)
    B = 16
    ls = pkg.LinearSystem(s, bank_size=B)
    data = None
    try:
        data = ls.collect(stride=1, max_samples=4)
        assert data.size == 4
        X, q = data.samples()
        vals, rows, cols = orc.assemble(s)
        m = orc.Mas(s, B)
        m.build(vals, rows, cols, 0, s.pairs)
        parts = mpnn.MasParts(m)
        A = mpnn.bcoo_to_csr(vals, rows, cols, s.n)
        graph = mpnn.coarse_graph(A, parts.map1, parts.cnt[1])
        assert data.nodes == graph.n
        res = _numpy_pcg_residuals(A, s.rhs, lambda r: parts.apply(r)[0], 4)
        for k in range(4):
            Xr, qr, _ = mpnn.sample(A, graph, parts, res[k])
            err = np.linalg.norm(X[k] - Xr) / np.linalg.norm(Xr)
            assert err <= 1e-5, (k, err)
            np.testing.assert_allclose(q[k], X[k][:, 2], rtol=0, atol=0)
        m.close()
    finally:
        if data is not None:
            data.close()  # before its context
        ls.close()


@pytest.mark.parametrize("coarse", [0, 1, 2])
def test_learned_level_on_a_held_out_scene(pkg, coarse):
    """Train on residuals of one seeded scene, solve another: the flexible
    solve with the correction reaches the tolerance, and the automatic mode
    ends on the variant with fewer iterations.  Prints what the network gains
    (or loses) against the same preconditioner without it."""
    B = 16
    train = pkg.LinearSystem(_scene(pkg, 1), bank_size=B, coarse_mode=coarse)
    test = pkg.LinearSystem(_scene(pkg, 2), bank_size=B, coarse_mode=coarse)
    data = net = None
    try:
        data = train.collect(stride=1, max_samples=200)
        for k in range(2):
            b = np.random.default_rng(40 + k).standard_normal(3 * train.scene.n)
            more = train.collect(stride=1, max_samples=200, b=b)
            X, q = more.samples()
            data.add(X, q)
            more.close()
        net = pkg.learned.Mpnn(test.ctx, layers=3, linear=False, seed=1)
        hist = net.train(data, epochs=12, batch=16, lr=1e-3, seed=0)
        assert np.all(np.isfinite(hist)) and hist[-1] < hist[0]
        plain = test.solve(tol=1e-5)
        test.set_learned(net, mode=1)
        learned = test.solve(tol=1e-5)
        assert learned.learned and not plain.learned
        assert learned.status == 0 and learned.true_rel_residual <= 1e-5
        print(f"coarse={coarse}: {data.size} samples, loss {hist[0]:.3e} -> {hist[-1]:.3e}; "
              f"held-out iterations: plain {plain.iterations}, learned {learned.iterations} "
              f"(gain {plain.iterations - learned.iterations:+d})")
        # automatic: one solve each way, then the better one
        test.set_learned(net, mode=2)
        runs = [test.solve(tol=1e-5) for _ in range(3)]
        assert [r.learned for r in runs[:2]] == [False, True]
        better_learned = runs[1].iterations < runs[0].iterations
        assert runs[2].learned == better_learned
        assert all(r.true_rel_residual <= 1e-5 for r in runs)
    finally:
        # the network and the training set use their contexts: close them first
        if net is not None:
            net.close()
        if data is not None:
            data.close()
        train.close()
        test.close()


def test_automatic_mode_drops_a_harmful_correction(pkg):
    """A network whose output is scaled by 1e4 spoils the flexible run: in
    automatic mode that run is capped at twice the plain count, the call then
    solves again without the correction (still converged), and the following
    solves stay plain."""
    ls = pkg.LinearSystem(_scene(pkg, 2), bank_size=16, coarse_mode=1)
    net = None
    try:
        net = pkg.learned.Mpnn(ls.ctx, layers=2, linear=True, seed=3)
        p = net.get_params()
        p[-64:] *= 1e4  # the output weights
        net.set_params(p)
        ls.set_learned(net, mode=2)
        plain = ls.solve(tol=1e-5)
        probe = ls.solve(tol=1e-5)
        assert not plain.learned
        assert probe.status == 0 and probe.true_rel_residual <= 1e-5
        if probe.learned:
            assert probe.iterations <= 2 * plain.iterations + 16
        else:
            assert probe.learned_abandoned
            # the capped run and the plain one both count
            assert probe.iterations >= plain.iterations
        after = [ls.solve(tol=1e-5) for _ in range(2)]
        assert all(not r.learned and r.status == 0 for r in after)
    finally:
        if net is not None:
            net.close()  # before its context
        ls.close()
